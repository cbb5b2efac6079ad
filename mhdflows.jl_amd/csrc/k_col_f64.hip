#define MHD_T double
#include "k_col.inc"
