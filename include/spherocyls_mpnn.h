/*
 * spherocyls_mpnn.h -- message-passing network over the contact graph that
 * proposes minimiser warm starts (no reference counterpart; the reference's
 * minimisers start from the unrelaxed state: gradient_descent_minimize,
 * src/cpp/relaxation_dynamics.cpp:33-85).
 *
 * Network: node input (cos 2phi, sin 2phi, R, A); edge input per contact
 * (closest-point distance d < R_i + R_j) (1 - d/sigma, contact normal x, y,
 * branch vector x, y); encoder to width 64; three message-passing layers
 *   m_ij = silu(W_m [h_i, h_j, e_ij] + b_m),  h_i += silu(W_u [h_i, sum_j m_ij] + b_u);
 * decoder to (dx, dy, dphi) per particle, multiplied by the output scale.
 * GEMM operands are BF16 on the MFMA units with FP32 accumulation; master
 * weights, gradients and Adam moments are FP32.
 *
 * Parameter vector layout (FP32, sc_mpnn_num_params entries): W_e [16][64],
 * then per layer W_m [144][64] and W_u [144][64], then W_o [80][16]; every
 * block is row-major [input][output] and its last real input row is the bias
 * (inputs are [x(4), 1, 0 x 11], [h_i, h_j, e(5), 1, 0 x 10], [h, M, 1, 0 x 15],
 * [h, 1, 0 x 15]).
 *
 * Graphs are given as CSR over receivers: offsets[n+1]; senders src[e] of
 * receiver i at offsets[i] .. offsets[i+1]-1, increasing; every edge i<-j needs
 * its reverse j<-i.  node_feat[n][4], edge_feat[e][5].
 */
#ifndef SPHEROCYLS_MPNN_H
#define SPHEROCYLS_MPNN_H

#include "spherocyls.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sc_mpnn sc_mpnn;

/* Create with seeded initial weights on `device`. */
int sc_mpnn_create(sc_mpnn** out, int device, unsigned long long seed);
int sc_mpnn_destroy(sc_mpnn* m);
const char* sc_mpnn_last_error(const sc_mpnn* m);
int sc_mpnn_num_params(const sc_mpnn* m, int* count);
int sc_mpnn_get_params(sc_mpnn* m, float* out);
int sc_mpnn_set_params(sc_mpnn* m, const float* in);
/* Gradients of the last sc_mpnn_backward. */
int sc_mpnn_get_grads(sc_mpnn* m, float* out);
/* Displacement = scale * network output (per component dx, dy, dphi). */
int sc_mpnn_set_output_scale(sc_mpnn* m, const float* scale3);
int sc_mpnn_get_output_scale(const sc_mpnn* m, float* scale3);

/* Load a graph from host arrays (training batches). */
int sc_mpnn_load_graph(sc_mpnn* m, int n, const float* node_feat, int e, const int* offsets,
                       const int* src, const float* edge_feat);
/* Forward pass on the loaded graph; out[n][3] (unscaled outputs) may be NULL. */
int sc_mpnn_forward(sc_mpnn* m, float* out);
/* Mean squared error against target[n][3] (unscaled) and its gradient with
 * respect to every parameter (after sc_mpnn_forward). */
int sc_mpnn_backward(sc_mpnn* m, const float* target, float* loss);
/* One Adam step on the gradients of the last backward pass (step counts from 1). */
int sc_mpnn_adam(sc_mpnn* m, float lr, float beta1, float beta2, float eps, int step);

/* Build the contact graph of a system's current state on the device (all
 * packings; node and edge counts returned). */
int sc_mpnn_graph_from_system(sc_mpnn* m, sc_system* s, int* n, int* e);
/* Copy the graph built by sc_mpnn_graph_from_system to host arrays, in the
 * system's caller particle order (offsets[n+1], src[e], node_feat[n][4],
 * edge_feat[e][5]; src and edge_feat may be NULL). */
int sc_mpnn_export_graph(sc_mpnn* m, sc_system* s, float* node_feat, int* offsets, int* src,
                         float* edge_feat);
/* Warm start: graph of the system's current state, forward pass, and every
 * particle displaced by scale * output (x, y in sheared coordinates, phi),
 * each component clamped to +-1 (x, y) and +-0.5 (phi); non-finite outputs
 * give no displacement.  Contacts of crossed rods (d = 0) carry a zero
 * normal. */
int sc_mpnn_warm_start(sc_mpnn* m, sc_system* s);
/* Damped warm start, all on the device: the proposal of sc_mpnn_warm_start is
 * applied scaled by each factor lambdas[0..nlam) in turn and every packing's
 * energy evaluated; each packing keeps the factor (or none, factor 0) whose
 * state has the lowest energy, so no packing starts above its unmoved energy.
 * e_before / e_after / lam_out [npack] (host, may be NULL): energy of the
 * unmoved state, energy kept, factor kept. */
int sc_mpnn_warm_start_damped(sc_mpnn* m, sc_system* s, int nlam, const double* lambdas,
                              double* e_before, double* e_after, double* lam_out);

#ifdef __cplusplus
}
#endif

#endif
