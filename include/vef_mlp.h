/* vef_mlp.h — C ABI of the per-node multilayer perceptron used as a learned
 * predictor in the arclength continuation.
 *
 * The reference predicts the next continuation state by the tangent or secant,
 * x_pred = x + t*dS (updateSolutionArclength and setDirectionVectors,
 * src/Continuation_Arclength_mod.F90:151-278,339-377).  This network has no
 * counterpart in the reference: it learns the per-node error of that tangent
 * prediction from the steps already converged and corrects the prediction;
 * the FP64 Newton iteration judges the corrected state, and the tangent
 * predictor remains the fallback (ve_flow_around_cylinder_fortran_amd/learned.py).
 *
 * Network: n_in inputs (<= 31; input 31 is held at 1 to carry the first bias)
 * -> 128 -> 128 -> 128 -> n_out outputs (<= 16), GELU (tanh form) after each
 * hidden layer.  Weights are kept in FP32 (Adam's state too) and used in BF16
 * by the matrix cores (16x16x16 BF16 MFMA, FP32 accumulation); activations and
 * back-propagated errors enter the matrix cores in BF16 as well.
 *
 * All arrays are device arrays, row-major [n][n_in] and [n][n_out], FP32; every
 * call is ordered on the given HIP stream.  Functions return 0 on success. */
#ifndef VEF_MLP_H
#define VEF_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vef_mlp vef_mlp;

enum { VEF_MLP_IN_MAX = 31, VEF_MLP_WIDTH = 128, VEF_MLP_OUT_MAX = 16 };

/* Creates a network with weights drawn from N(0, 1/fan_in) (seeded) and zero
 * biases. */
int vef_mlp_create(int32_t n_in, int32_t n_out, uint64_t seed, vef_mlp** out);

/* Parameters as one FP32 vector (host): W1[32][128] (row 31 = first bias),
 * W2[128][128], b2[128], W3[128][128], b3[128], W4[128][16], b4[16]; weights
 * stored [in][out]. */
int64_t vef_mlp_num_params(const vef_mlp* m);
int vef_mlp_get_params(const vef_mlp* m, float* host_params);
int vef_mlp_set_params(vef_mlp* m, const float* host_params);

/* Y = network(X) for n rows. */
int vef_mlp_forward(vef_mlp* m, int64_t n, const float* X, float* Y, void* hip_stream);

/* Gradient of the loss mean((Y - T)^2) over the n rows and the n_out outputs
 * with respect to the parameters (device FP32 vector in the layout above), and
 * the loss (device float[1]); either output may be NULL. */
int vef_mlp_gradient(vef_mlp* m, int64_t n, const float* X, const float* T, float* grad, float* loss,
                     void* hip_stream);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on that loss. */
int vef_mlp_adam_step(vef_mlp* m, int64_t n, const float* X, const float* T, float lr, float* loss,
                      void* hip_stream);

const char* vef_mlp_last_error(const vef_mlp* m);
void vef_mlp_destroy(vef_mlp* m);

#ifdef __cplusplus
}
#endif

#endif
