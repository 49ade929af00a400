// page kernels for 32-bit elements
#define FLS_W 32
#include "k_page_w.inc"
