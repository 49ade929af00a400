// page kernels for 64-bit elements
#define FLS_W 64
#include "k_page_w.inc"
