// Exact (IEEE, no FMA contraction) stencil: build with -ffp-contract=off.
#define FIP_FAST_TU 0
#include "stencil_impl.inc"
