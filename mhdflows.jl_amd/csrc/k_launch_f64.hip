#define MHD_T double
#include "k_launch.inc"
