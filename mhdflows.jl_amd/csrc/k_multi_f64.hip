#define MHD_T double
#include "k_multi.inc"
