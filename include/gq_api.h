/*
 * gq_api.h -- C-ABI of the MI355X-native Fisher-Kolmogorov forward solve and
 * likelihood path (libgq.so).
 *
 * Plain C types only: opaque context, plain pointers and sizes, status codes.
 * One host thread per context; every call is ordered on the context's HIP
 * stream and returns when its results are on the host (or resident in device
 * memory for calls that return nothing).  Device memory is owned by the
 * context; input buffers stay owned by the caller and are copied.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the reference's source/ tree).
 */
#ifndef GQ_API_H
#define GQ_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    GQ_OK = 0,
    GQ_ERR_ARG = 1,         /* invalid argument (sizes, ids, ranges) */
    GQ_ERR_HIP = 2,         /* a HIP runtime call failed */
    GQ_ERR_NOMEM = 3,       /* device or host allocation failed */
    GQ_ERR_STATE = 4,       /* call out of order (no grid, no observations) */
    GQ_ERR_TOPOLOGY = 5,    /* grid invariant violated (level jump > 1) */
    GQ_ERR_UNSUPPORTED = 6, /* configuration outside what the path supports */
    GQ_ERR_NUMERIC = 7      /* non-finite parameter or time step */
} gq_status;

typedef struct gq_ctx gq_ctx;

/* MRAG::BlockInfo (MRAGcore/MRAGCommon.h:166-260) for a 16^3 block on the
 * unit cube: origin = index * 2^-level (vertex-centred), h = 2^-level / 16. */
typedef struct {
    int32_t block_id; /* BlockInfo::blockID (stable until refine/compress) */
    int16_t index[3]; /* BlockInfo::index */
    int16_t level;    /* BlockInfo::level */
    float origin[3];  /* BlockInfo::origin */
    float h;          /* BlockInfo::h[0] (isotropic) */
} gq_block_info;

/* Parameters of Glioma_ReactionDiffusionOperator(Dw, Dg, rho) and
 * UpdateTumor(dt) (Glioma/Glioma_ReactionDiffusionOperator.h:20,232), in
 * unit-cube units: Dw and Dg are already divided by L^2
 * (Glioma/Glioma_HG_UQ.cpp:532-533). */
typedef struct {
    double Dw, Dg, rho, dt;
} gq_fk_params;

enum { GQ_FIELD_PHI = 0, GQ_FIELD_PWM = 1, GQ_FIELD_PGM = 2 };

/* ---- context ---------------------------------------------------------- */
gq_status gq_ctx_create(gq_ctx **out, int device);
gq_status gq_ctx_destroy(gq_ctx *ctx);
const char *gq_status_string(gq_status s);
/* Message of the last failing call on this context ("" if none). */
const char *gq_last_error(const gq_ctx *ctx);

/* ---- grid: replaces Grid<W,B>(bpd,bpd,bpd,maxStencil) with
 * Refiner_SpaceExtension(resJump=1, maxLevel) and Compressor(1)
 * (Glioma/Glioma_HG_UQ.cpp:28-34, MRAGcore/MRAGrid.inl:57-111).
 * blocks_per_dim must be a power of two; wavelet_order is 2
 * (Wavelets_Interp2ndOrder, the reference default, MRAGHeaders.h:14) or 4
 * (Wavelets_Interp4thOrder).  Block size is 16^3. */
gq_status gq_grid_create(gq_ctx *ctx, int blocks_per_dim, int max_level, int wavelet_order);
/* Number of leaf blocks (Grid::getBlocksInfo().size()). */
gq_status gq_grid_block_count(gq_ctx *ctx, int *n);
/* Grid::getBlocksInfo (MRAGcore/MRAGrid.inl:191-219).  The order returned
 * here is the order of every per-block array of this API. */
gq_status gq_grid_blocks_info(gq_ctx *ctx, gq_block_info *out, int capacity, int *n);
/* Block data transfer: n_blocks x 4096 floats, block order of
 * gq_grid_blocks_info, x fastest inside a block (Block::operator()(ix,iy,iz)). */
gq_status gq_grid_write_field(gq_ctx *ctx, int field, const float *data, int n_blocks);
gq_status gq_grid_read_field(gq_ctx *ctx, int field, float *data, int n_blocks);

/* Point-sample a uniform gpd^3 volume (x fastest, gpd = 16*blocks_per_dim)
 * onto the leaves: a leaf at level l takes every 2^(log2(blocks_per_dim)-l)-th
 * point (the reference's _ic_* initialisers evaluate each field at the block
 * points, Glioma/Glioma_HG_UQ.cpp:67-225).  sample = -1 writes every sample
 * of an ensemble. */
gq_status gq_grid_write_volume(gq_ctx *ctx, int field, const float *vol, int gpd, int sample);

/* ---- ensembles (no reference counterpart: the reference runs one solve per
 * process, one process per sampler proposal, Glioma/main.cpp).  nsamples
 * independent grids of one configuration share the context's block pool;
 * gq_grid_create is the one-sample case.  Single-grid calls (blocks_info,
 * read/write, fwt, refine/compress, dump, likelihoods) address the selected
 * sample (default 0); gq_fk_steps advances every sample with one parameter
 * set. */
gq_status gq_ensemble_create(gq_ctx *ctx, int nsamples, int blocks_per_dim, int max_level, int wavelet_order);
/* Every sample back to the uniform base-level grid of the configuration it
 * was created with, keeping the device block pool (no allocation where it
 * fits, as gq_ensemble_create would do): the way to start a new batch of
 * adaptive solves.  Field contents are undefined until written (anatomy by
 * gq_grid_write_volume, phi by it or gq_ensemble_ic_gaussian). */
gq_status gq_ensemble_reset(gq_ctx *ctx);
gq_status gq_ensemble_size(gq_ctx *ctx, int *nsamples);
gq_status gq_sample_select(gq_ctx *ctx, int sample);
/* Leaf-block count of every sample (nsamples int32). */
gq_status gq_ensemble_block_counts(gq_ctx *ctx, int32_t *out);

enum { GQ_RUN_GRAPH = 1 }; /* replay each chunk of steps as one HIP graph */
typedef struct {
    int32_t adapt_every;  /* steps between adaptations (0: no adaptation) */
    int32_t max_level;    /* refinement limit */
    double tol_refine;    /* Science::AutomaticRefinement tolerance */
    double tol_compress;  /* Science::AutomaticCompression tolerance */
    int32_t flags;        /* GQ_RUN_* */
} gq_run_opts;
/* Glioma_HG_UQ::run (Glioma/Glioma_HG_UQ.cpp:511-588) for every sample at
 * once: sample s takes n_steps[s] steps with p[s]; after every full chunk of
 * adapt_every steps it runs AutomaticRefinement(tol_refine, max_level, 1
 * loop) + AutomaticCompression(tol_compress), and at the end one more
 * refinement (:574-575).  Each step is one launch of the step kernel over the
 * blocks of all samples still running; the result of every sample equals
 * that of a one-sample grid run alone on the same schedule. */
gq_status gq_ensemble_run(gq_ctx *ctx, const gq_fk_params *p, const int32_t *n_steps, const gq_run_opts *opts);
/* Statistics of the last gq_ensemble_run: {cell updates, device ms of the
 * step chunks, chunks, adaptation rounds, host seconds in adaptation, final
 * work-list blocks, last slab depth, graph flag}; after gq_adjoint_flip:
 * {forward cell updates, forward ms, backward ms, segments, 0, blocks, slab
 * depth, 0}. */
gq_status gq_ensemble_stats(gq_ctx *ctx, double out[8]);
/* gq_loglik_flip for every sample: out = nsamples x {logL, L_T1, L_FLAIR, L_PET}. */
gq_status gq_ensemble_loglik_flip(gq_ctx *ctx, const double params[4], double *out);

/* ---- discrete adjoint (no reference counterpart: the reference samples
 * without gradients).  Uniform grids only (every leaf at the base level).
 *
 * Gaussian initial condition per sample, ic[s*5] = {x0, y0, z0, A, sigma}
 * (voxel units of the gpd^3 grid, periodic distance):
 * u0 = A exp(-|x - x0|^2 / (2 sigma^2)). */
gq_status gq_ensemble_ic_gaussian(gq_ctx *ctx, const double *ic);
/* For every sample: n_steps forward steps with p[s] (Dw, Dg > 0) from the
 * current state (or from the Gaussian initial condition ic, if not null),
 * keeping a checkpoint every ckpt_every steps; the flip-rate log-likelihood
 * at the final state (lik = {th_T1, th_FLAIR, flip_rate, pet_sigma}, the
 * observations of gq_obs_flip_upload), out_ll[s*4] as gq_loglik_flip; and the
 * exact derivative of the discrete FP32 forward map, with recomputation from
 * the checkpoints: out_grad[s*6] = dL/d{log Dw (Dg = Dw/10 scaled with it),
 * log rho, x0, y0, z0, A} (the IC terms are 0 when ic is null).  The
 * Bernoulli terms are piecewise constant in u: their derivative is zero
 * almost everywhere and the gradient is that of the PET term.  The final
 * state is left in the grids. */
gq_status gq_adjoint_flip(gq_ctx *ctx, const gq_fk_params *p, int n_steps, int ckpt_every, const double lik[4],
                          const double *ic, double *out_ll, double *out_grad);

/* ---- learned surrogate (no reference counterpart): a 3D convolutional
 * encoder-decoder on 64^3 that maps {4 parameters broadcast as channels,
 * the initial condition u0, pw, pg} to the final tumour field, trained and
 * evaluated by bf16 MFMA kernels with FP32 accumulation (FP32 master
 * weights, Adam).  Fields are 64^3 floats, x fastest; params are 4 floats
 * per item.  Loss: mean (p - t)^2 + 0.1 mean BCE(sigmoid(20 (p - 0.25)),
 * [t > 0.25]).  One context per surrogate; batch <= the capacity given at
 * creation (<= 64). */
typedef struct gq_sur gq_sur;
gq_status gq_sur_create(gq_sur **out, int device, int batch_capacity, uint64_t seed);
gq_status gq_sur_destroy(gq_sur *s);
const char *gq_sur_last_error(const gq_sur *s);
gq_status gq_sur_set_anatomy(gq_sur *s, const float *pw, const float *pg);
gq_status gq_sur_param_count(gq_sur *s, int64_t *n);
/* all parameters: per conv (E1..E4, D3, D2, D1) weights [cout][27*cin] then
 * bias [cout]; head weights [16] and bias */
gq_status gq_sur_get_params(gq_sur *s, float *out);
gq_status gq_sur_set_params(gq_sur *s, const float *in);
gq_status gq_sur_predict(gq_sur *s, const float *params, const float *u0, int batch, float *out);
/* loss and its gradient w.r.t. every parameter (gq_sur_get_params layout), no update */
gq_status gq_sur_loss_grad(gq_sur *s, const float *params, const float *u0, const float *target, int batch,
                           double *loss, float *grads);
/* one Adam step (lr) on a batch; loss before the step */
gq_status gq_sur_train_step(gq_sur *s, const float *params, const float *u0, const float *target, int batch, float lr,
                            double *loss);
/* device ms of the last forward and backward pass */
gq_status gq_sur_timing(gq_sur *s, double out[2]);
/* flip-rate likelihood of the surrogate's predictions (the first stage of
 * delayed-acceptance sampling): observations on the 64^3 grid as
 * gq_obs_flip_upload; out = batch x {logL, L_T1, L_FLAIR, L_PET}, FP64 sums */
gq_status gq_sur_obs_flip_upload(gq_sur *s, const uint8_t *t1, const uint8_t *t2, const float *pet);
gq_status gq_sur_loglik_flip(gq_sur *s, const float *params, const float *u0, int batch, const double lik[4],
                             double *out);

/* ---- forward solve: replaces Glioma_HG_UQ::_reactionDiffusionStep
 * (Glioma/Glioma_HG_UQ.cpp:266-277) = BlockProcessing::pipeline_process(vInfo,
 * collection, boundaryInfo, Glioma_ReactionDiffusionOperator) followed by
 * BlockProcessing::process(vInfo, collection, UpdateTumor)
 * (MRAGmultithreading/MRAGBlockProcessing_TBB.h:476-512, :249-264), fused,
 * n_steps times with a global dt. */
gq_status gq_fk_steps(gq_ctx *ctx, const gq_fk_params *p, int n_steps);

/* ---- adaptation ---------------------------------------------------------
 * BlockFWT detail per block, FWTReport::getOverAll_DetailMaxMag with the
 * phi projector (MRAGcore/MRAGBlockFWT.h:49-192,220-298;
 * Glioma/Glioma_Types.h:405-413).  FP64; out has one double per block in
 * gq_grid_blocks_info order. */
gq_status gq_fwt_details(gq_ctx *ctx, double *out, int n_blocks);
/* Grid::refine(set<int>) / Grid::compress(set<int>, int&) on explicit block
 * ids (MRAGcore/MRAGrid.inl:248-355, :384-486), with the split/collapse data
 * transfer of BlockSplitter/BlockCollapser (MRAGBlockSplitter.h:24-99,
 * MRAGBlockCollapser.h:24-85).  n_out = parents split / blocks collapsed. */
gq_status gq_refine_blocks(gq_ctx *ctx, const int32_t *block_ids, int n, int *n_out);
gq_status gq_compress_blocks(gq_ctx *ctx, const int32_t *block_ids, int n, int *n_out);
/* Science::AutomaticRefinement<0,0>(grid, fwt, tol, max_level, max_loops)
 * and Science::AutomaticCompression<0,0>(grid, fwt, tol, max_loops)
 * (MRAGscience/MRAGScienceCore.h:42-108, :117-177). */
gq_status gq_auto_refine(gq_ctx *ctx, double tol, int max_level, int max_loops, int *n_out);
gq_status gq_auto_compress(gq_ctx *ctx, double tol, int max_loops, int *n_out);

/* ---- output for the likelihood: Glioma_HG_UQ::_dumpUQoutput
 * (Glioma/Glioma_HG_UQ.cpp:309-372): phi on the uniform gpd^3 grid,
 * gpd = blocks_per_dim*16, piecewise-constant replication.  out: gpd^3
 * doubles, x fastest, to be zero-initialised by the caller (voxels the
 * reference leaves unwritten keep their value). */
gq_status gq_dump_uniform(gq_ctx *ctx, double *out, size_t n_out);

/* ---- likelihood -----------------------------------------------------------
 * Reference model: HGG_Likelihood::run (Likelihood/source/HGG_Likelihood.cpp:
 * 40-189).  Observations on the gpd^3 grid (x fastest) and, for
 * selectivePoints, the Sphere.dat integer points (npts x 3, xyz per row).
 * params = {PETsigma2, PETscale, T1uc, T2uc, slope} (LikelihoodInput.txt).
 * The model field is the current grid, mapped as gq_dump_uniform maps it.
 * out = {logL, L_pet, L_T1, L_T2}; FP64 accumulation on the device. */
gq_status gq_obs_ref_upload(gq_ctx *ctx, int gpd, const double *pet, const double *t1, const double *t2,
                            const int32_t *points_xyz, int npts);
gq_status gq_loglik_ref(gq_ctx *ctx, const double params[5], int selective, int step_pet, int step_ti,
                        double out[4]);

/* The same reference model on an explicit model volume (the HGG_data.dat
 * contract of the reference's likelihood program, Likelihood_main.cpp:18-36):
 * model is gpd^3 doubles, x fastest; observations from gq_obs_ref_upload. */
gq_status gq_loglik_ref_volume(gq_ctx *ctx, const double *model, int gpd, const double params[5], int selective,
                               int step_pet, int step_ti, double out[4]);

/* Flip-rate Bernoulli + Gaussian PET model (no reference counterpart: the
 * likelihood of the synthetic-observation workloads).  t1/t2: 0/1 masks,
 * pet: value or NaN where not observed, gpd^3 voxels x fastest.  The
 * observation grid may be coarser than the model's by an integer factor s
 * (observation voxel o is compared with model voxel s*o).
 * params = {th_T1, th_FLAIR, flip_rate, pet_sigma}; out = {logL, L_T1,
 * L_FLAIR, L_PET}. FP64 accumulation on the device. */
gq_status gq_obs_flip_upload(gq_ctx *ctx, int gpd, const uint8_t *t1, const uint8_t *t2, const float *pet);
gq_status gq_loglik_flip(gq_ctx *ctx, const double params[4], double out[4]);

/* ---- no allocation ---------------------------------------------------------
 * No reference counterpart: for callers that cannot allocate (code under HIP
 * graph capture, tight sampler loops, harnesses that forbid hipMalloc).  The
 * flip-rate log-likelihood of gq_loglik_flip on a model volume already in
 * device memory, with the same kernels and the same bits: model is gpd^3
 * doubles on the device (x fastest); t1, t2, pet are device arrays on an
 * obs_gpd^3 grid (obs_gpd divides gpd; observation voxel o is compared with
 * model voxel (gpd/obs_gpd)*o); out is 4 device doubles {logL, L_T1,
 * L_FLAIR, L_PET}; stream is a hipStream_t (null: the default stream).  The
 * call only enqueues kernels on the stream and returns: it allocates nothing
 * and does not synchronise.  Its partial sums live in a fixed device array
 * of the library (48 KiB per GPU, kept until the process ends): calls that
 * overlap on one GPU race on it.  The current HIP device must be the one that
 * holds the arrays and the stream (the array's per-GPU copy is looked up on
 * the current device).  gpd is at most GQ_NOALLOC_MAX_GPD. */
#define GQ_NOALLOC_MAX_GPD 1024
gq_status gq_loglik_flip_volume_noalloc(const double *model, int gpd, const uint8_t *t1, const uint8_t *t2,
                                        const float *pet, int obs_gpd, const double params[4], double *out,
                                        void *stream);

/* ---- measurement ----------------------------------------------------------
 * Device time of the last gq_fk_steps call, summed over its timed step-kernel
 * launches (ms), with the number of those launches and the leaf cells they
 * updated (every launch and cell of the call without per-kernel timing). */
gq_status gq_last_step_timing(gq_ctx *ctx, double *step_kernel_ms, int *launches, int64_t *cell_updates);
/* per_kernel = k >= 1: time every k-th step-kernel launch of gq_fk_steps
 * (counted across calls from this call on; 1: every launch) with a pair of
 * HIP events that the kernel's dispatch records at its start and end
 * (hipExtLaunchKernel: the events cost host time per launch, so a stride
 * keeps the timed solve close to an untimed one); 0 (default): one pair
 * around the whole call, which then includes the cross-level ghost kernels. */
gq_status gq_set_step_timing(gq_ctx *ctx, int per_kernel);
/* Device time of every step-kernel launch timed since the last drain (with
 * per-kernel timing) or of every chunk of steps (without), in stream order:
 * ms[i] and the cell updates cells[i] it covers.  Stepping does not wait for
 * its timing events; this call does.  ms == NULL: *n = the number of
 * entries, which are kept; else cap must be at least that number
 * (GQ_ERR_ARG otherwise, entries kept), and the entries are copied and
 * dropped. */
gq_status gq_step_timing_drain(gq_ctx *ctx, float *ms, int64_t *cells, int cap, int *n);
/* Step autotuner.  mode 1 (default): time the slab depths 16/8/4/2 (planes
 * of a 16^3 block per workgroup) with the cross-level face ghosts from a
 * separate launch, on the grid the first time a block count of a new
 * power-of-two size is stepped, and keep the fastest; mode 4: also time
 * depth 1 (one workgroup per block streaming its planes through registers)
 * and every depth with the face ghosts evaluated inside the step; mode 0: a
 * fixed size rule; mode 2: always fixed_zt (16/8/4/2/1) with the separate
 * launch; mode 3: always fixed_zt with the in-step evaluation.  Every
 * configuration gives the same bits. */
gq_status gq_set_autotune(gq_ctx *ctx, int mode, int fixed_zt);
/* Step coefficients (no reference counterpart; the reference's functor
 * recomputes D/b from p_w, p_g per face, Glioma_ReactionDiffusionOperator.h:
 * 174-188).  precomputed = 1 (default): 1/K of every cell is computed once
 * per anatomy, topology and parameter set, and a step reads phi and 1/K
 * (8 B per cell) instead of phi, pw, pg (12 B); 0: the step recomputes 1/K
 * from pw, pg every step.  Both give the same bits. */
gq_status gq_set_step_coeffs(gq_ctx *ctx, int precomputed);
/* Step fusion (no reference counterpart; the reference applies one step per
 * pipeline_process/process pair, Glioma_HG_UQ.cpp:266-277).  on = 1: two
 * consecutive steps over a work list without cross-level faces (a uniform
 * grid) run as one kernel that stages phi and 1/K with a two-cell halo and
 * recomputes the neighbours' first-step halo values, so HBM sees one read and
 * one write per two steps; 0 (default): one launch per step, which measured
 * faster on MI355X (the fused kernel is bound by its redundant first-step
 * arithmetic).  Both give the same bits.  Needs the precomputed step
 * coefficients. */
gq_status gq_set_step_fusion(gq_ctx *ctx, int on);
/* Persistent step chunks (no reference counterpart; the reference runs one
 * pipeline_process/process pair per step, Glioma_HG_UQ.cpp:266-277).  A chunk
 * of up to 16 steps between adaptations runs as one cooperative launch whose
 * workgroups stay resident, with grid-wide barriers between the face-ghost
 * and step phases of every step.  mode 0 (default): never -- measured
 * slower on MI355X, where each grid-wide barrier costs more than a kernel
 * launch; 1: an autotuner candidate, timed at the fastest slab depth beside
 * one launch per step and kept where faster; 2: always (error where the
 * device cannot launch it cooperatively; slab depths 16-2).  Every mode
 * gives the same bits. */
gq_status gq_set_step_persistent(gq_ctx *ctx, int mode);
/* Whether the last step call ran persistent chunks, and the autotuner's last
 * timing of the persistent candidate in us per step (0 if not timed). */
gq_status gq_persistent_info(gq_ctx *ctx, int *used, double *tuned_us);
/* Slab depth and face-ghost mode (1 = evaluated in the step) used by the last
 * step call, and the last autotune timings in us per step: [0..4] depths 16,
 * 8, 4, 2, 1 with the separate face-ghost launch (both kernels timed), [5..9]
 * the same depths with the in-step evaluation (0 where not timed). */
gq_status gq_autotune_info(gq_ctx *ctx, int *zt, int *fused, double us[10]);
/* Host-side seconds spent in {ghost-table builds, refine/compress planning,
 * FWT calls, splits, collapses, autotuning} and the counts of table builds
 * and FWT calls, accumulated over the context's life. */
gq_status gq_host_stats(gq_ctx *ctx, double out[8]);
/* on = 1 (default): gq_fk_steps captures its launches as a HIP graph and
 * replays it (relaunching a cached graph when the launches repeat), unless
 * every launch is timed (gq_set_step_timing 1); 0: direct launches.  Both give
 * the same bits (no reference counterpart). */
gq_status gq_set_step_graphs(gq_ctx *ctx, int on);
/* Graph-captured step chunks of gq_fk_steps and gq_ensemble_run (no reference counterpart):
 * chunks captured and instantiated, chunks relaunched from an instantiated
 * graph with identical launches, and graphs currently kept (at most 8, the
 * least recently used one is destroyed first; graphs are never updated in
 * place). */
gq_status gq_graph_info(gq_ctx *ctx, int64_t *captures, int64_t *relaunches, int *cached);
/* Device-memory footprint of the context (bytes) and pool capacity (blocks). */
gq_status gq_memory_info(gq_ctx *ctx, int64_t *bytes, int *pool_blocks);

/* ---- host-only topology (no device): the Grid's refine/compress topology
 * and ghost-point expansion without block data, for callers that plan on
 * the host and for checking the planner without a GPU.  Leaves are reported
 * in the same (Morton) order as gq_grid_blocks_info. */
typedef struct gq_topo gq_topo;
gq_status gq_topo_create(gq_topo **out, int blocks_per_dim, int max_level, int wavelet_order);
gq_status gq_topo_destroy(gq_topo *t);
/* level[n], index[3n], block_id[n] in leaf order; n = number of leaves. */
gq_status gq_topo_leaves(gq_topo *t, int32_t *level, int32_t *index, int32_t *block_id, int capacity, int *n);
/* Grid::refine / Grid::compress topology update for explicit block ids. */
gq_status gq_topo_refine(gq_topo *t, const int32_t *block_ids, int n, int *n_out);
gq_status gq_topo_compress(gq_topo *t, const int32_t *block_ids, int n, int *n_out);
/* Ghost value of leaf `leaf` (leaf order) at local coordinates loc (may lie
 * outside [0,16)): merged (leaf, cell, weight) terms in canonical order. */
gq_status gq_topo_point_terms(gq_topo *t, int leaf, const int32_t loc[3], int capacity, int32_t *out_leaf,
                              int32_t *out_cell, double *out_w, int *n);

#ifdef __cplusplus
}
#endif

#endif /* GQ_API_H */
