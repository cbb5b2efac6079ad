#define MHD_T float
#include "k_launch.inc"
