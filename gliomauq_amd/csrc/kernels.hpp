// kernels.hpp -- argument blocks and launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gq {

struct StepArgs {
    const float *phi_in;
    float *phi_out;
    const float *pw;
    const float *pg;
    const int *work;   // Morton-ordered slots
    int nwork;
    const int *fnbr;   // [slot*6+f]
    const int *fbase;  // [slot*6+f]
    const float *fik;  // [slot*6+f][256]
    const float *gphi; // cross-level face ghosts of phi_in
    const int *level;  // [slot]
    const float *ih2_lvl;  // 1/h^2 per level
    float Dw, Dg, rho, dt;
};

struct LabArgs {
    const int *nbr26;
    const int *rbase26;
    const int *ptr;
    const int *src;
    const double *w;
    const int *row_begin;  // [slot] lab rows of a slot: [row_begin, row_end)
    const int *row_end;
};

struct UniformArgs {
    const double *vol;  // if set: a uniform gpd^3 model volume (x fastest) replaces the grid view
    const float *phi;
    const int *lut_slot;
    const int *level;
    const int *index;  // [slot*3]
    int gpd;
    double hole;
};

// zt: slab depth (16, 8, 4 or 2 planes per workgroup)
hipError_t launch_fk_step(const StepArgs &a, int zt, hipStream_t st);
hipError_t launch_face_ghosts(const float *phi, const int *ptr, const int *src, const float *w, float *out, int nrows,
                              hipStream_t st);
hipError_t launch_face_prep(const float *pw, const float *pg, const int *work, int nwork, const int *fnbr,
                            const int *fbase, const int *ptr, const int *src, const float *w, float *fik, float Dw,
                            float Dg, hipStream_t st);
// Evaluate the cross-level lab rows of the listed slots for one field (FP64,
// canonical row order) into ghost[row].
hipError_t launch_lab_ghosts(const LabArgs &a, const float *fld, const int *slots, int nslots, double *ghost,
                             hipStream_t st);
hipError_t launch_fwt_detail(int order, const LabArgs &a, const double *ghost, const float *phi, const int *cand,
                             int ncand, double *out, hipStream_t st);
// ghost3: three ghost arrays (phi, pw, pg) of nrows each
hipError_t launch_split(int order, const LabArgs &a, const double *ghost3, int nrows, float *phi, float *pw,
                        float *pg, const int *parents, const int *children, int nparents, hipStream_t st);
hipError_t launch_collapse(float *phi, float *pw, float *pg, const int *parents, const int *children, int nparents,
                           hipStream_t st);
hipError_t launch_dump_uniform(const UniformArgs &u, double *out, hipStream_t st);
int reduce_blocks(int64_t n);
hipError_t launch_lik_flip(const UniformArgs &u, const uint8_t *t1, const uint8_t *t2, const float *pet, double th1,
                           double th2, double lp1, double lp0, double c0, double is2, double *partial, int nblocks,
                           double *out, hipStream_t st);
hipError_t launch_lik_ref(const UniformArgs &u, const double *pet, const double *t1, const double *t2,
                          const int *pts, int npts, int selective, int step_pet, int step_ti, double scale,
                          double t1uc, double t2uc, double slope, double *partial, int nblocks, double *out,
                          hipStream_t st);

}  // namespace gq

namespace gq {
hipError_t launch_scatter_blocks(float *pool, const float *dense, const int *slots, int n, hipStream_t st);
hipError_t launch_gather_blocks(const float *pool, float *dense, const int *slots, int n, hipStream_t st);
}  // namespace gq
