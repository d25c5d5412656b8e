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

// Location tables for ghost evaluation on the device: the leaf covering a
// point is found through a look-up over the finest level.
struct LabArgs {
    const int *nbr26;  // [slot*26] same-level neighbour slot per region or -1
    const int *lut;    // (2^lut_level)^3 finest block cells -> slot
    const int *level;  // [slot]
    const int *index;  // [slot*3]
    int lut_level;
    int order;         // 2 or 4
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
// Ghost values of the cross-level points of the wide lab of each listed slot
// (FP64, depth-first order) into ghost[k * lab_points + q].
// for_split: the points the split reads; else the points the FWT reads.
hipError_t launch_lab_ghosts(const LabArgs &a, const float *fld, const int *slots, int nslots, double *ghost,
                             bool for_split, hipStream_t st);
int lab_points(int order);  // points of the wide lab (Interp4: 21^3, Interp2: 18^3)
hipError_t launch_fwt_detail(int order, const LabArgs &a, const double *ghost, const float *phi, const int *cand,
                             int ncand, double *out, hipStream_t st);
// ghost3: ghost arrays of phi, pw, pg, each nparents * lab_points
hipError_t launch_split(int order, const LabArgs &a, const double *ghost3, float *phi, float *pw, float *pg,
                        const int *parents, const int *children, int nparents, hipStream_t st);
// Step rows of the cross-level faces: faces[k] = slot*6+f; rows k*256 + u + 16v.
// Pass 1 counts the depth-first terms of every row, pass 2 writes them.
hipError_t launch_face_rows_count(const LabArgs &a, const int *faces, int nfaces, int *counts, hipStream_t st);
hipError_t launch_face_rows_fill(const LabArgs &a, const int *faces, int nfaces, const int *ptr, int *src, float *w,
                                 hipStream_t st);
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
