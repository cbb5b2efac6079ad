#define MHD_T float
#include "k_misc.inc"
