// page kernels for 16-bit elements: UNCOMPRESSED, FFOR and DELTA pages
#define FLS_W 16
#define FLS_PART 1
#include "k_page_w.inc"
