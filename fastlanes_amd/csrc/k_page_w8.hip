// page kernels for 8-bit elements
#define FLS_W 8
#include "k_page_w.inc"
