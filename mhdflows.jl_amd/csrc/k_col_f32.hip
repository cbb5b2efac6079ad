#define MHD_T float
#include "k_col.inc"
