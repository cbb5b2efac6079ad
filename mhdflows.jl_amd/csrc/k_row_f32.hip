#define MHD_T float
#include "k_row.inc"
