// kernels.hpp -- argument blocks and launchers of the HIP kernels.
#pragma once
#include <vector>

#include <hip/hip_runtime.h>

#include <cstdint>

#include "grid.hpp"  // BS, look-up entry layout

namespace gq {

struct StepArgs {
    const float *phi_in;
    float *phi_out;
    const float *pw;
    const float *pg;
    const int *work;   // Morton-ordered slots
    int nwork;
    const int *fnbr;   // [slot*6+f]
    const float *fik;  // [slot*6+f][256]
    const float *ikc;  // [slot*4096] per-cell 1/K codes (k_face_prep), or null: 1/K from pw, pg every step
    const float *gphi; // cross-level face ghosts of phi_in, or null: the step evaluates the rows itself
    const int64_t *fptr;  // face-ghost rows (ptr/src/w) for the in-kernel evaluation
    const int *fsrc;
    const float *fw;
    const int *level;  // [slot]
    const float *ih2_lvl;  // 1/h^2 per level
    float Dw, Dg, rho, dt;     // used when sprm is null
    const int *slot_sample;    // [slot] sample of the block (ensembles)
    const float *sprm;         // [sample*4] Dw, Dg, rho, dt (ensembles) or null
    // [w*16] per work item: slot, 1/h^2 (float bits), fnbr[6], absolute first face-ghost row
    // of each face (-1 for same-level faces), sample, 0 -- one scalar load
    // gives a workgroup everything the slot tables would give it through dependent look-ups
    const int *wrec;
    const int *nbr26;  // [slot*26] same-level neighbour per region (the fused two-step kernel)
};
constexpr int WREC = 16;

// One cooperative launch over m <= PMAX steps (k_fk_persistent): step j
// updates the first nwork[j] work items and evaluates the first nrows[j]
// cross-level face-ghost rows of phi[c] into gout (a.gphi == gout; with
// a.gphi null the step evaluates the rows itself); c = c0 ^ (j & 1).
constexpr int PMAX = 16;
struct PersistArgs {
    StepArgs a;
    float *phi[2];
    float *gout;
    int c0, m;
    int nwork[PMAX];
    int nrows[PMAX];
};

// Location tables for ghost evaluation on the device: the leaf covering a
// point is found through a look-up over the finest level.
struct LabArgs {
    const int *nbr26;  // [slot*26] same-level neighbour slot per region or -1
    // per sample (2^lut_level)^3 finest block cells -> slot | level << LUT_LEVEL_SHIFT, concatenated
    const int *lut;
    const int *lutoff; // [slot] offset of the slot's sample look-up in lut
    const int *level;  // [slot]
    const int *index;  // [slot*3]
    int lut_level;
    int order;         // 2 or 4
};

struct UniformArgs {
    const double *vol;  // if set: a uniform gpd^3 model volume (x fastest) replaces the grid view
    const float *phi;
    const int *lut_slot;  // bpd^3 per sample, concatenated (sample = blockIdx.y)
    const int *level;
    const int *index;  // [slot*3]
    int gpd;
    double hole;
    int obs_stride;  // flip likelihood: observation voxel o reads model voxel o*obs_stride (obs grid gpd/obs_stride)
};

// One backward step of the discrete adjoint on uniform grids (every face
// neighbour at the same level): lambda_n from lambda_{n+1}, u_n and u_{n+1}.
struct AdjArgs {
    const float *u;       // u_n
    const float *ik;      // 1/K per cell (launch_adj_prep)
    const float *mu_in;   // mu_{n+1} = [0 < u_{n+1} < 1] lambda_{n+1}
    float *out;           // mu_n, or lambda_n when out_lambda
    int out_lambda;
    const int *work;
    int nwork;
    const int *fnbr;
    const int *level;
    const float *ih2_lvl;
    const int *slot_sample;
    const float *sprm;    // [sample*4] Dw, Dg, rho, dt
    double *gacc;         // [(slot*adj_slabs()+slab)*2]: dL/dlogDw, dL/dlogrho
};
int adj_slabs();
hipError_t launch_adj_step(const AdjArgs &a, hipStream_t st);
// ik = 1/K of every cell of the listed slots for the slot's sample's (Dw, Dg)
hipError_t launch_adj_prep(const float *pw, const float *pg, const int *work, int nwork, const int *slot_sample,
                           const float *sprm, float *ik, hipStream_t st);
// mu = [0 < u < 1] dL/du of the flip-rate model's PET term at u (uniform grids; observations on the grid
// coarser by stride)
hipError_t launch_lik_flip_grad(const float *u, const float *pet, const int *work, int nwork, const int *index,
                                int gpd, int stride, double inv_s2, float *mu, hipStream_t st);
// Gaussian initial condition per sample, icp[s*5] = x0, y0, z0 (voxels), A, sigma (voxels)
hipError_t launch_ic_gaussian(float *phi, const int *work, int nwork, const int *index, const int *slot_sample,
                              const double *icp, int gpd, hipStream_t st);
// out[k*4 + {0..3}] = dL/d(x0, y0, z0, A) of work slot k from lambda_0
hipError_t launch_ic_grad(const float *lam0, const int *work, int nwork, const int *index, const int *slot_sample,
                          const double *icp, int gpd, double *out, hipStream_t st);

// zt: slab depth (16, 8, 4 or 2 planes per workgroup), or 1: one workgroup per
// block streaming its planes through registers (k_fk_step_col)
// Two fused steps phi_in -> phi_out (distinct buffers) for a work list
// whose blocks have all 26 neighbours at their own level (no cross-level
// faces); needs the per-cell 1/K codes.
hipError_t launch_fk_step2(const StepArgs &a, hipStream_t st, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// e0/e1 (optional): events the dispatch itself records at the kernel's start and end
hipError_t launch_fk_step(const StepArgs &a, int zt, hipStream_t st, hipEvent_t e0 = nullptr,
                          hipEvent_t e1 = nullptr);
hipError_t launch_face_ghosts(const float *phi, const int64_t *ptr, const int *src, const float *w, float *out,
                              int nrows, hipStream_t st);
// m steps in one cooperative launch at slab depth zt (16/8/4/2);
// hipErrorNotSupported where the device cannot launch it cooperatively.
hipError_t launch_fk_persistent(const PersistArgs &p, int zt, hipStream_t st);
// sprm: per-sample Dw, Dg (ensembles) or null for (Dw, Dg)
hipError_t launch_face_prep(const float *pw, const float *pg, const int *work, int nwork, const int *fnbr,
                            const int *fbase, const int *sfbase, const int *slot_sample, const int64_t *ptr,
                            const int *src, const float *w, float *fik, float *ikc, float Dw, float Dg,
                            const float *sprm, hipStream_t st);
// Slot-indexed tables and the scatter of packed rows {slot, level, index[3],
// sample, lutoff, fnbr[6], fbase[6], nbr26[26]} into them.
struct SlotTables {
    int *level, *index, *slot_sample, *lutoff, *fnbr, *fbase, *nbr26;
};
hipError_t launch_scatter_slot_rows(const SlotTables &t, const int *packed, int nrows, hipStream_t st);
// Ghost values of the cross-level points of the wide lab of each listed slot
// (FP64, depth-first order) into ghost[k * lab_points + q].
// for_split: the points the split reads; else the points the FWT reads.
// nf = 1 or 3 fields evaluated in one walk; field f goes to ghost + f * field_stride.
struct FieldPtrs {
    const float *p[3];
};
// Cross-level lab points of the listed slots: items = j << 8 | region << 3 |
// chunk (region 0..25: the regions around the block in (dx, dy, dz) order,
// centre left out; chunks of 64 points) for the regions whose neighbour is
// not a same-level leaf (lab_items builds them).
hipError_t launch_lab_ghosts(const LabArgs &a, const float *const *flds, int nf, const int *slots,
                             const int *items, int nitems, double *ghost, size_t field_stride, bool for_split,
                             hipStream_t st);
// Host: the items of launch_lab_ghosts for `slots` from the slot-indexed
// same-level neighbour table nbr26 (regions without used points skipped).
void lab_items(const std::vector<int> &slots, const std::vector<int> &nbr26, int order, std::vector<int> &items);
int lab_points(int order);  // points of the wide lab (Interp4: 21^3, Interp2: 18^3)
hipError_t launch_fwt_detail(int order, const LabArgs &a, const double *ghost, const float *phi, const int *cand,
                             int ncand, double *out, hipStream_t st);
// ghost3: ghost arrays of phi, pw, pg, each nparents * lab_points
hipError_t launch_split(int order, const LabArgs &a, const double *ghost3, float *phi, float *pw, float *pg,
                        const int *parents, const int *children, int nparents, hipStream_t st);
// Step rows of the cross-level faces: faces[k] = slot*6+f; rows k*256 + u + 16v.
// Pass 1 counts the depth-first terms of every row, pass 2 writes them.
hipError_t launch_face_rows_count(const LabArgs &a, const int *faces, int nfaces, int *counts, hipStream_t st);
hipError_t launch_face_rows_fill(const LabArgs &a, const int *faces, int nfaces, const int64_t *ptr, int *src,
                                 float *w, hipStream_t st);
// ptr[0] = 0, ptr[1..n] = inclusive sum of counts (64-bit).  With temp null
// only sets temp_bytes.
hipError_t scan_counts(const int *counts, int64_t *ptr, int n, void *temp, size_t &temp_bytes, hipStream_t st);
hipError_t launch_collapse(float *phi, float *pw, float *pg, const int *parents, const int *children, int nparents,
                           hipStream_t st);
hipError_t launch_dump_uniform(const UniformArgs &u, double *out, hipStream_t st);
int reduce_blocks(int64_t n);
// workgroups of the flip-rate likelihood for n_obs observation voxels (at most 2048)
int flip_blocks(int64_t n_obs);
// nsamples views (lut_slot per sample); partial: nsamples*nblocks*3, out: nsamples*3
hipError_t launch_lik_flip(const UniformArgs &u, int nsamples, const uint8_t *t1, const uint8_t *t2, const float *pet,
                           double th1, double th2, double lp1, double lp0, double c0, double is2, double *partial,
                           int nblocks, double *out, hipStream_t st);
hipError_t launch_lik_ref(const UniformArgs &u, const double *pet, const double *t1, const double *t2,
                          const int *pts, int npts, int selective, int step_pet, int step_ti, double scale,
                          double t1uc, double t2uc, double slope, double *partial, int nblocks, double *out,
                          hipStream_t st);

}  // namespace gq

namespace gq {
hipError_t launch_scatter_blocks(float *pool, const float *dense, const int *slots, int n, hipStream_t st);
hipError_t launch_gather_blocks(const float *pool, float *dense, const int *slots, int n, hipStream_t st);
hipError_t launch_copy_blocks(float *dst, const float *src, const int *slots, int n, hipStream_t st);
// Point-sample a uniform gpd^3 volume (gpd = 16 * 2^base_level) onto the listed slots.
hipError_t launch_sample_volume(float *pool, const float *vol, int gpd, int base_level, const int *slots, int n,
                                const int *level, const int *index, hipStream_t st);
}  // namespace gq
