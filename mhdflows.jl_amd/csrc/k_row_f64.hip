#define MHD_T double
#include "k_row.inc"
