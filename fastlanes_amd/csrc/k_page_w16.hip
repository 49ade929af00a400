// page kernels for 16-bit elements
#define FLS_W 16
#include "k_page_w.inc"
