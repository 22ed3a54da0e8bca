/* Learned edge prior for the CVRP ant colony: a 12 -> 64 -> 64 -> 1 ReLU MLP
 * scoring candidate edges (i, nn[i][r]), BF16 weights on the matrix cores with
 * FP32 accumulation, trained with weighted binary cross-entropy on edges of
 * best-so-far tours of the plain colony.  Its output multiplies the heuristic
 * eta (aco_problem.prior takes p^beta, see aco_prior_build).
 *
 * There is no counterpart in the reference (cvrp_ocl computes eta from the
 * distance alone, kernel.cl:672-679,782-789); every function here is the
 * build's own.  All pointers are host pointers; device memory is owned by the
 * library.  Return codes are those of cvrp_aco.h (ACO_OK = 0, negative on
 * error, aco_last_error() for the message).
 */
#ifndef CVRP_PRIOR_H
#define CVRP_PRIOR_H

#include <stdint.h>

#include "cvrp_aco.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ACO_PRIOR_FEATURES 12
#define ACO_PRIOR_HIDDEN   64
/* parameter vector (fp32): W1[64][12], b1[64], W2[64][64], b2[64], w3[64], b3 */
#define ACO_PRIOR_PARAMS   (64 * 12 + 64 + 64 * 64 + 64 + 64 + 1)

typedef struct aco_prior aco_prior;

typedef struct aco_prior_train_params {
    int32_t  epochs;
    int32_t  batch;        /* edges per Adam step (<= 65536)                  */
    float    lr;           /* Adam step size                                  */
    float    lr_final;     /* cosine decay from lr to lr_final over the run   */
    float    pos_weight;   /* loss weight of positive (tour) edges            */
    float    beta1, beta2, eps;
    uint64_t shuffle_seed; /* SplitMix64 Fisher-Yates order of each epoch     */
} aco_prior_train_params;

void aco_prior_default_train_params(aco_prior_train_params *tp);

/* Glorot-uniform weights from a SplitMix64 stream seeded with `seed`, zero
 * biases, zero Adam moments. */
int aco_prior_create(uint64_t seed, aco_prior **out);
int aco_prior_destroy(aco_prior *net);
int aco_prior_get_params(aco_prior *net, float *params, int32_t count);
/* replaces the weights and resets the Adam moments and step counter */
int aco_prior_set_params(aco_prior *net, const float *params, int32_t count);

/* Features of the n*k candidate edges (i, nn[i][r]), r < k, i-major, 12 per
 * edge: d(i,j)/mkd, r/k, dem_i/Q, dem_j/Q, angle_i/pi, angle_j/pi, rad_i/R,
 * rad_j/R, kd_i/mkd, kd_j/mkd, (d(0,i)+d(0,j)-d(i,j))/R, [i in nn[j][:k]];
 * polar coordinates around node 0 (the depot), R the largest radius, kd_v
 * the distance to v's k-th candidate, mkd its mean.  nn is n x nn_width
 * (k <= nn_width).  Computed on the GPU. */
int aco_prior_features(int32_t n, const float *x, const float *y, const int32_t *demand,
                       int32_t capacity, const float *dist, const int32_t *nn,
                       int32_t nn_width, int32_t k, float *feat);

/* Logits of `count` edges (features count x 12). */
int aco_prior_forward(aco_prior *net, const float *feat, int64_t count, float *logits);

/* Gradient of L = mean_e w_e (softplus(z_e) - y_e z_e), w_e = pos_weight for
 * y_e = 1 else 1, over `count` edges (<= 65536), without updating. */
int aco_prior_gradient(aco_prior *net, const float *feat, const float *label, int64_t count,
                       float pos_weight, float *grad, float *loss);

/* One Adam step with the given gradient (step counter advances). */
int aco_prior_adam_step(aco_prior *net, const float *grad, float lr, float beta1, float beta2,
                        float eps);

/* Minibatch training on the GPU: per epoch a Fisher-Yates permutation, per
 * batch one fused forward/backward launch and one reduce + Adam launch.
 * epoch_loss (optional, epochs entries) receives the mean batch loss. */
int aco_prior_train(aco_prior *net, const float *feat, const float *label, int64_t count,
                    const aco_prior_train_params *tp, float *epoch_loss);

/* Dense n x n multiplier for aco_problem.prior: for candidate edges
 * (i, nn[i][r]), r < k, sigmoid(logit)^exponent; floor^exponent elsewhere.
 * With exponent = beta the learned probability multiplies eta. */
int aco_prior_build(aco_prior *net, int32_t n, const float *x, const float *y,
                    const int32_t *demand, int32_t capacity, const float *dist,
                    const int32_t *nn, int32_t nn_width, int32_t k, float exponent, float floor,
                    float *prior);

/* The same multiplier applied on the device to every instance of a context
 * created without a prior (before aco_init): features from the context's
 * own distances, demands and candidate lists (the construction list when
 * k <= params.k, else the 2-opt neighbour list), the MLP, then
 * heur[i][j] *= prior[i][j].  x[b], y[b] are instance b's coordinates.
 * Equal to creating the context with aco_problem.prior = aco_prior_build(). */
int aco_prior_apply(aco_ctx *ctx, aco_prior *net, const float *const *x, const float *const *y,
                    int32_t k, float exponent, float floor);

#ifdef __cplusplus
}
#endif

#endif /* CVRP_PRIOR_H */
