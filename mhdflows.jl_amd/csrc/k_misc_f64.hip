#define MHD_T double
#include "k_misc.inc"
