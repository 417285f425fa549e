// Learned preconditioner (xg_gnn.hip) as seen by the CG driver (xg_cg.hip).
#pragma once
#include <cstdint>

#include "xg_ctx.hpp"

// number of network parameters (flat FP32 layout: we[32] be[32], per layer
// W[32x64] b[32], wo[32] bo[1])
#define XG_GNN_NPARAM 6337

// Ahat and sqrt(d) for the matrix (ia, ja, a) with preconditioner diagonal d
int xg_gnn_prepare_internal(xg_ctx* c, const int* ia, const int* ja, const double* a,
                            const double* d, int64_t n);
// z = network(r), then the Jacobi fallback where (z, r) <= 0
int xg_gnn_apply_internal(xg_ctx* c, const int* ia, const int* ja, const double* r,
                          const double* d, double* z);
// whether the last application fell back to Jacobi (synchronises)
int xg_gnn_fallback_flag(xg_ctx* c, int* used);
// z = network(r) keeping the activations for a backward pass
int xg_gnn_forward_internal(xg_ctx* c, const int* ia, const int* ja, const double* r,
                            const double* d, double* z);
// gradients of the pair loss added to the accumulated gradient; *loss if asked
int xg_gnn_backward_internal(xg_ctx* c, const int* ia, const int* ja, const double* a,
                             const double* r, double etr, const double* z, double* loss);
double* xg_gnn_zbuf(xg_ctx* c);
void xg_gnn_free(xg_ctx* c);
