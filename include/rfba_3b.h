/*
 * rfba_3b.h -- C ABI of the MI355X (gfx950) 3B likelihood library
 * (librfba_3b.so): the per-block spot log-likelihood of RFBA's 3B stage,
 * marginalised over on/off/bleached blink states, with its gradient.
 *
 * Model (reference): spot shape and its derivative
 *   RFBA Main/storm.h:51-80 (spot_shape_s, spot_shape_diff_position),
 * per-frame log-probabilities of a spot over a background
 *   RFBA Main/spot_with_background.hh:106-151,
 * HMM forward recursion in log space with ln(0) = -1e100
 *   RFBA Main/forward_algorithm.h:34-40, :157-256, :344-382,
 * and the conjugate-gradient functor over sampled backgrounds
 *   RFBA Main/multispot5.cc:579-669.
 * The reference has no C ABI for this path (an in-process C++ functor); the
 * entry points below are what its callers (ConjugateGradientOnly<5>::
 * find_next_point, conjugate_gradient_only.h:106-232; GibbsSampler2::next,
 * sampled_multispot.h:681-709) would bind.
 *
 * Precision: spot images and residuals in FP32; per-frame sums, the
 * forward/backward recursions and all reductions in FP64.  The recursions
 * keep every message value with an exponent of its own (extended range),
 * so a state whose weight falls below the double range keeps it, as in
 * the reference's log space.  The gradient is the adjoint (reverse mode):
 * forward messages are kept only every `checkpoint` frames and recomputed
 * segment by segment during the backward pass.
 *
 * Conventions: parameters phi = (I, sigma, x, y, z) (storm.h:71); states
 * 0 = on, 1 = off, 2 = bleached; pixel coordinates are integer voxel
 * positions (sampled_multispot.h:170-177); frames x pixels arrays are
 * frame-major [T][P].  Return 0 on success, negative on error
 * (rfba3b_last_error).
 */
#ifndef RFBA_3B_H
#define RFBA_3B_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rfba3b_model {
    double A[9];          /* A[3*i + j] = P(state i -> state j) (multispot5.cfg:15) */
    double pi[3];         /* initial state probabilities (multispot5.cfg:16) */
    double variance;      /* Gaussian pixel-noise variance (multispot5.cc:1440) */
    double zfactor;       /* z scale of the spot shape (storm.h:55) */
    int z_exact;          /* 0: the reference's d/dz (storm.h:78, no 1/zF^2); 1: exact */
    int checkpoint;       /* forward-message checkpoint stride; 0 = ceil(sqrt(T)) */
    double mu_brightness, sigma_brightness;  /* log-normal prior of I (multispot5.cc:660-661) */
    double mu_blur, sigma_blur;              /* log-normal prior of sigma (:662) */
} rfba3b_model;

/* Reference functor, SpotNegProbabilityDiffWithSampledBackground::operator()
 * (multispot5.cc:616-669), for n candidate parameter vectors of one spot:
 *   g = -( (1/S) sum_s d/dphi logL(phi; bg_s) + [dlnLN(I), dlnLN(sigma), 0, 0, 0] )
 * NaN rows for I <= 0 or sigma <= 0 (multispot5.cc:618-619).
 * Host buffers: xyz [P][3] pixel coordinates, data [T][P], bg [S][T][P]
 * (sampled intensities of the other spots), phi [n][5].
 * Outputs: grad [n][5]; logL [n][S] may be NULL. */
int rfba3b_neg_grad_sampled(const int *xyz, int P, int T, const float *data, const float *bg, int S,
                            const double *phi, int n, const rfba3b_model *model, double *grad,
                            double *logL);

/* Resident form of the same functor, created once per optimised spot as the
 * reference builds SampledBackgroundData (multispot5.cc:538-575) and then
 * called by ConjugateGradientOnly<5> at least twice per iteration
 * (conjugate_gradient_only.h:92, 136, 180, 207).  Host inputs, as the
 * reference's: xyz [P][3] (pixels), data [T][P] (pixel_intensities) and
 * bg [S][T][P] (sample_intensities_without_spot), all FP64.  Creation forms
 * each residual data - bg_s in FP64, rounds it once to FP32 and keeps it on
 * the device with its per-frame FP64 sum of squares; max_candidates bounds
 * n of every evaluation (<= 4096, as T and S).  The handle owns its device
 * buffers. */
typedef struct rfba3b_functor rfba3b_functor;  /* opaque handle */
int rfba3b_functor_create(const int *xyz, int P, int T, const double *data, const double *bg, int S,
                          const rfba3b_model *model, int max_candidates, rfba3b_functor **handle);
/* operator()(phi) for n <= max_candidates candidate vectors phi [n][5]:
 * grad [n][5] as rfba3b_neg_grad_sampled (NaN rows for I <= 0 or
 * sigma <= 0), logL [n][S] (may be NULL).  Allocates nothing: two kernel
 * launches on `stream` and copies through the handle's pinned staging
 * buffer; returns when the results are on the host.  Calls on one handle
 * are serialised. */
int rfba3b_functor_eval(rfba3b_functor *handle, const double *phi, int n, double *grad, double *logL,
                        void *stream);
/* sampled_background_spot_hessian_ffbs (multispot5.cc:902-982), the
 * Hessian the model-selection step decomposes (multispot5.cc:1981, 2074),
 * for the handle's data at spot phi [5]: per background sample the spot's
 * forward messages (forward_algorithm_delta), bs_iterations backward
 * samplings (draw k of sampling (s, it) is uniforms[(s*bs + it)*T + k], the
 * last frame first, as rfba3b_ffbs), and over each sampled path the sums of
 * the per-frame parts sum_p e H - d d^T and sum_p e d of the on frames
 * (spot_shape_hess_position, storm.h:86-150); hess [5][5] (row-major) gets
 * sum (H + D D^T) / (bs S var) - mean(D) mean(D)^T plus the log-normal
 * priors' second derivatives on [0][0] and [1][1].  deriv [5] (may be NULL)
 * gets mean(D) plus the priors' derivatives; paths [S][bs][T] (may be NULL)
 * the sampled states.  With z_exact = 0 the spot Hessian keeps the
 * reference's z terms and takes its never-set entry [4][3] (storm.h:143)
 * as 0; z_exact = 1 uses the exact second derivatives.  The first call, and
 * a call with a larger S x bs_iterations, allocates the handle's Hessian
 * buffers.  Error for I <= 0 or sigma <= 0. */
int rfba3b_functor_hessian_ffbs(rfba3b_functor *handle, const double *phi, int bs_iterations,
                                const double *uniforms, double *hess, double *deriv, uint8_t *paths,
                                void *stream);
void rfba3b_functor_destroy(rfba3b_functor *handle);

/* Forward messages delta_t(i) (forward_algorithm.h:344-382) of one spot over
 * one background, for forward filtering / backward sampling.
 * Host buffers; delta [T][3]; returns logL in *logL (may be NULL). */
int rfba3b_forward_messages(const int *xyz, int P, int T, const float *data, const float *bg,
                            const double *phi, const rfba3b_model *model, double *delta, double *logL);

/* ---- batched model evaluation over blocks -------------------------------
 * A block holds a dense X x Y x Z voxel box with T frames of data, a
 * constant background b0, K spots, Q parameter sets for all K spots and S
 * sampled state paths of every spot.  Work item (q, k, s): spot k at
 * params[q][k] over the background b0 + sum_{j != k} [state_s,j(t) == on]
 * mu(params[q][j]) (the other spots' sampled intensities, as remove_spot in
 * multispot5.cc:1602-1603), giving logL[q][k][s] and grad[q][k][s][5]. */
/* Forward filtering / backward sampling of one spot's blink states: the
 * forward messages of rfba3b_forward_messages, then backward sampling
 * (forward_algorithm.h:541-566) on the device.  uniforms[T] are the
 * generator's draws in the reference's order (uniforms[0] picks the last
 * frame's state).  states[T] gets 0 on / 1 off / 2 bleached; delta [T][3]
 * and logL may be NULL.  Host buffers. */
int rfba3b_ffbs(const int *xyz, int P, int T, const float *data, const float *bg,
                const double *phi, const rfba3b_model *model,
                const double *uniforms, uint8_t *states, double *delta,
                double *logL);

/* Backward sampling of n chains at once on the device: d_delta [n][T][3]
 * forward messages, d_uniforms [n][T] draws, d_states [n][T] output.
 * Asynchronous on stream. */
int rfba3b_backward_sample_device(const double *d_delta, int n, int T,
                                  const rfba3b_model *model,
                                  const double *d_uniforms, uint8_t *d_states,
                                  void *stream);

/* The reference's random stream (MT19937 as seeded by mt19937.h:48-51 and
 * mersenne.cpp RandomInit; draws as mt19937.h:61-63): n draws after skipping
 * `skip` draws, for callers that reproduce a reference run. */
int rfba3b_mt19937_uniforms(int seed, unsigned long long skip, size_t n, double *out);

typedef struct rfba3b_block {
    int dims[3];              /* X, Y, Z (pixel (x,y,z) has coordinates x, y, z) */
    int T, K, Q, S;
    double b0;                /* constant background */
    const float *d_data;      /* device [T][Z][Y][X] */
    const double *d_params;   /* device [Q][K][5] */
    const uint8_t *d_states;  /* device [S][K][T], 0 = on */
    double *d_logL;           /* device [Q][K][S] */
    double *d_grad;           /* device [Q][K][S][5], NULL for logL only */
} rfba3b_block;

typedef struct rfba3b_batch rfba3b_batch;  /* opaque plan */

int rfba3b_batch_create(const rfba3b_block *blocks, int n, const rfba3b_model *model, rfba3b_batch **plan);
/* All blocks at once: per memory chunk, three launches (residual planes,
 * spot correlations, HMM forward + adjoint) over every block, parameter set
 * and background sample; optionally captured into a hipGraph. */
int rfba3b_batch_run(rfba3b_batch *plan, void *stream, int use_graph, float *kernel_ms);
/* Plain form for comparison, same kernels and results: one launch sequence
 * per block (per_param_set = 0) or per (block, parameter set) (1). */
int rfba3b_batch_run_per_block(rfba3b_batch *plan, void *stream, int per_param_set, float *kernel_ms);
/* Number of memory chunks the plan's residual planes are split into (each
 * chunk's planes fit the plane budget: RFBA3B_PLANE_BUDGET_MB in the
 * environment at rfba3b_batch_create, default 16384 MiB). */
int rfba3b_batch_chunks(const rfba3b_batch *plan);
/* HMM form of the plan's runs (forward_algorithm_deriv,
 * forward_algorithm.h:157-332): 0 (default) chooses by the longest series,
 * one wave per item for T <= 512 (each lane holds its frames' messages in
 * registers; nothing is stored) and one lane per item above; 1 always runs
 * the lane form, which keeps the forward messages every `checkpoint` frames
 * of the model (ceil(sqrt(T)) for 0) and recomputes the rest in the adjoint.
 * Returns the previous form, or a negative status. */
int rfba3b_batch_set_hmm_form(rfba3b_batch *plan, int form);
void rfba3b_batch_destroy(rfba3b_batch *plan);

/* ---- Gibbs sampling of the blink states --------------------------------
 * The reference's GibbsSampler2 (sampled_multispot.h:362-730), whose next()
 * (:681-709) visits every spot in turn: remove the spot from the current
 * sample intensities over its mask (remove_spot, :115-122), compute the
 * masked on/off log-probabilities of every frame (SpotWithBackgroundMasked,
 * spot_with_background.hh:106-151), run the forward recursion
 * (forward_algorithm_delta2, forward_algorithm.h:394-433), draw the spot's
 * state path by backward sampling (:542-566) and add the spot back
 * (add_spot, sampled_multispot.h:124-131).
 *
 * A chain is one such sampler over a list of P pixels with K spots.  Spot
 * k's masked on/off log-probabilities differ, in frame t, by
 * -d_t(k) / (2 variance) with
 *   d_t(k) = S2_k - 2 (C_t(k) - A_t(k)),
 *   S2_k = sum over M_k of spot_k^2,  C_t(k) = sum over M_k of spot_k (data_t - base),
 *   A_t(k) = sum over the other spots j on in frame t of O_kj,
 *   O_kj = sum over M_k and M_j of spot_k spot_j,
 * because the sample intensities hold each on spot over its own mask M_j
 * (add_spot); everything else is common to the three states of a frame.
 * The plan computes S2, C and O once, on the device, and keeps A in FP64
 * as states change, so a sweep reads no pixels.  base = 0 reproduces the
 * reference's empty sample.  Each chain owns an MT19937 generator seeded
 * like the reference's (mt19937.h:48-51, 37 draws discarded), advanced by
 * `skip` draws; it persists across runs.  One spot update consumes exactly
 * T draws (backward_sampling). */
typedef struct rfba3b_gibbs_chain {
    int P, T, K;
    const float *d_data;       /* device [T][P] pixel intensities */
    const double *d_spot;      /* device [K][P] spot intensities (compute_spot_intensity) */
    const int *d_mask_off;     /* device [K + 1]: spot k's mask is d_mask[off[k] .. off[k+1]) */
    const int *d_mask;         /* device pixel indices in [0, P) (spot_pixels) */
    double base;               /* intensity of the empty sample; 0 for the reference */
    uint8_t *d_sample;         /* device [K][T] current sample (in/out) */
    uint8_t *d_samples;        /* device [S][K][T]: the sample after each of the S steps
                                  of rfba3b_gibbs_run (the layout of rfba3b_block.d_states),
                                  or NULL */
    unsigned int seed;         /* MT19937 seed */
    unsigned long long skip;   /* draws skipped after seeding */
} rfba3b_gibbs_chain;

typedef struct rfba3b_gibbs rfba3b_gibbs;  /* opaque plan */

/* Validates the chains and their masks, computes S2, C and O, seeds the
 * generators; the caller's buffers stay caller-owned and must outlive the
 * plan.  Starts from reset.  setup_ms (may be NULL) gets the statistics'
 * kernel time.  Limits: at most 65535 chains, T <= 2836 frames per chain
 * (a chain's forward messages are held in on-chip memory), K <= 19456
 * spots; each mask strictly increasing within [0, P). */
int rfba3b_gibbs_create(const rfba3b_gibbs_chain *chains, int n, const rfba3b_model *model, rfba3b_gibbs **plan,
                        float *setup_ms);
/* GibbsSampler2::reset (sampled_multispot.h:430-438): every spot state 2
 * (off, bleached) in every frame, sample intensities = base.  Asynchronous. */
int rfba3b_gibbs_reset(rfba3b_gibbs *plan, void *stream);
/* `samples` consecutive next() calls of every chain (one launch, one
 * 64-thread workgroup per chain), each of `sample_iterations` sweeps over
 * the spots in index order; after step s the sample is copied to
 * d_samples[s].
 * variance: host [samples] noise variance of each step (set_variance,
 * :423-426), or NULL for the model's.  Synchronous; kernel_ms may be NULL. */
int rfba3b_gibbs_run(rfba3b_gibbs *plan, int samples, int sample_iterations, const double *variance, void *stream,
                     float *kernel_ms);
/* GibbsSampler2::sample_intensities (sampled_multispot.h:725-728) of chain
 * i, plus base, from its current sample: d_out [T][P].  Asynchronous. */
int rfba3b_gibbs_sample_intensities(const rfba3b_gibbs *plan, int chain, double *d_out, void *stream);
/* Draws consumed by chain i since its seeding, skip included. */
int rfba3b_gibbs_draws(const rfba3b_gibbs *plan, int chain, unsigned long long *draws);
void rfba3b_gibbs_destroy(rfba3b_gibbs *plan);

/* get_spot_pixels (multispot5.cc:1001-1020) for K spots: the pixels of a
 * ball of radius phi[1] * 6 + 1 (getBall, :1256-1269) around the spot's
 * rounded position (ir_rounded3D, :989-995), found in the pixel list xyz
 * [P][3] (sorted by z, then y, then x, as the reference's region).  Host
 * buffers.  offsets [K + 1] is always filled; idx (capacity cap) gets the
 * pixel indices when non-NULL. */
int rfba3b_spot_pixels(const int *xyz, int P, const double *phi, int K, int *offsets, int *idx, size_t cap);
/* compute_spot_intensity (sampled_multispot.h:183-191, storm.h:51-57,
 * 229-235) of K spots at P pixels on the device: d_xyz [P][3], d_phi [K][5],
 * d_spot [K][P] (FP64).  Asynchronous. */
int rfba3b_spot_intensities_device(const int *d_xyz, int P, const double *d_phi, int K, double zfactor,
                                   double *d_spot, void *stream);

const char *rfba3b_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* RFBA_3B_H */
