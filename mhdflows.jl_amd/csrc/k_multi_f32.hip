#define MHD_T float
#include "k_multi.inc"
