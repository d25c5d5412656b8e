// gq_mrag.hpp -- header-only C++ mirror of the reference's block-operator API
// over the C-ABI of gq_api.h.  A driver written against MRAG + the Glioma
// operators (Glioma/Glioma_HG_UQ.cpp) keeps its source:
//
//   Grid<W,B> grid(bpd,bpd,bpd)              -> gq::mrag::Grid<W,B> grid(ctx, bpd, maxLevel)
//   grid.getBlocksInfo(), getBlockCollection(), getBoundaryInfo()
//                                            -> same names
//   blockProcessing.pipeline_process(vInfo, collection, boundaryInfo, rhs)
//   BlockProcessing::process(vInfo, collection, updateTumor, granularity)
//                                            -> same calls; the pair runs one fused step
//   Science::AutomaticRefinement<0,0>(grid, blockfwt, tol, maxLevel, loops, &profiler)
//   Science::AutomaticCompression<0,0>(grid, blockfwt, tol, loops, &profiler)
//                                            -> same calls (BlockFWT, Profiler: stand-in types)
//   Glioma_HG_UQ::_dumpUQoutput(grid)        -> grid.dumpUniform()
//
// Block data live on the GPU; the collection and boundary-info objects are
// handles to the grid's device state, the FWT and profiler arguments carry no
// state.  Errors become gq::mrag::Error exceptions carrying the gq_status.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "gq_api.h"

namespace gq {
namespace mrag {

struct Error : std::runtime_error {
    gq_status status;
    Error(gq_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

inline void check(gq_ctx *c, gq_status s, const char *what)
{
    if (s != GQ_OK) throw Error(s, std::string(what) + ": " + gq_status_string(s) + " " + (c ? gq_last_error(c) : ""));
}

class Context {
public:
    explicit Context(int device = 0) { check(nullptr, gq_ctx_create(&c_, device), "gq_ctx_create"); }
    ~Context() { if (c_) gq_ctx_destroy(c_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    gq_ctx *get() const { return c_; }

private:
    gq_ctx *c_ = nullptr;
};

typedef gq_block_info BlockInfo;

// Wavelet and block types of the reference's Grid<W,B> (MRAGcore/
// MRAGWavelets_Interp2ndOrder.h, MRAGWavelets_Interp4thOrder.h,
// MRAGcore/MRAGBlock.h); only their order and size are used.
struct Wavelets_Interp2ndOrder {
    static const int order = 2;
};
struct Wavelets_Interp4thOrder {
    static const int order = 4;
};
template <typename Element, int sX = 16, int sY = sX, int sZ = sX>
struct Block {
    typedef Element ElementType;
    static const int sizeX = sX, sizeY = sY, sizeZ = sZ;
    static_assert(sX == 16 && sY == 16 && sZ == 16, "blocks are 16^3 on this path");
};
typedef Block<float, 16, 16, 16> Block16;

// BlockFWT<W, B, Projector> and Profiler of the reference's adaptation calls
// (MRAGcore/MRAGBlockFWT.h:195-478, MRAGcore/MRAGProfiler.h): stateless here,
// the transform runs on the device with the phi projector.
template <typename W, typename B, typename Projector = void>
struct BlockFWT {
};
struct Profiler {
};

// Glioma_ReactionDiffusionOperator<3>(Dw, Dg, rho): stencil [-1,2)
template <int DIM = 3>
struct Glioma_ReactionDiffusionOperator {
    static_assert(DIM == 3, "the path is 3D");
    int stencil_start[3] = {-1, -1, -1};
    int stencil_end[3] = {2, 2, 2};
    double Dw, Dg, rho;
    Glioma_ReactionDiffusionOperator(double Dw_, double Dg_, double rho_) : Dw(Dw_), Dg(Dg_), rho(rho_) {}
};

// UpdateTumor<3>(dt)
template <int DIM = 3>
struct UpdateTumor {
    static_assert(DIM == 3, "the path is 3D");
    double dt;
    explicit UpdateTumor(double dt_) : dt(dt_) {}
};

// Grid::getBlockCollection() (MRAGcore/MRAGrid.h:79): the block data, on the
// device; a handle naming the grid's context.
template <typename B>
class BlockCollection {
public:
    explicit BlockCollection(gq_ctx *c) : ctx_(c) {}
    gq_ctx *ctx() const { return ctx_; }

private:
    gq_ctx *ctx_;
};

// Grid::getBoundaryInfo() (MRAGcore/MRAGrid.h:80): the ghost information of
// the grid, evaluated on the device; a handle here.
struct BoundaryInfo {
    gq_ctx *ctx = nullptr;
};

template <typename W = Wavelets_Interp2ndOrder, typename B = Block16>
class Grid {
public:
    Grid(Context &ctx, int bpd, int maxLevel, int waveletOrder = W::order)
        : ctx_(ctx), bpd_(bpd), maxLevel_(maxLevel), coll_(ctx.get())
    {
        check(ctx.get(), gq_grid_create(ctx.get(), bpd, maxLevel, waveletOrder), "gq_grid_create");
        binfo_.ctx = ctx.get();
    }
    Grid(const Grid &) = delete;
    Grid &operator=(const Grid &) = delete;
    int size() const
    {
        int n = 0;
        check(ctx_.get(), gq_grid_block_count(ctx_.get(), &n), "gq_grid_block_count");
        return n;
    }
    std::vector<BlockInfo> getBlocksInfo() const
    {
        std::vector<BlockInfo> v((size_t)size());
        int n = 0;
        check(ctx_.get(), gq_grid_blocks_info(ctx_.get(), v.data(), (int)v.size(), &n), "gq_grid_blocks_info");
        v.resize((size_t)n);
        return v;
    }
    const BlockCollection<B> &getBlockCollection() const { return coll_; }
    BoundaryInfo &getBoundaryInfo() { return binfo_; }
    int getCurrentMaxLevel() const
    {
        int m = 0;
        for (const BlockInfo &b : getBlocksInfo()) m = b.level > m ? b.level : m;
        return m;
    }
    void write(int field, const std::vector<float> &blocks)
    {
        check(ctx_.get(), gq_grid_write_field(ctx_.get(), field, blocks.data(), (int)(blocks.size() / 4096)), "write");
    }
    std::vector<float> read(int field) const
    {
        std::vector<float> v((size_t)size() * 4096);
        check(ctx_.get(), gq_grid_read_field(ctx_.get(), field, v.data(), (int)(v.size() / 4096)), "read");
        return v;
    }
    int refine(const std::vector<int32_t> &ids)
    {
        int n = 0;
        check(ctx_.get(), gq_refine_blocks(ctx_.get(), ids.data(), (int)ids.size(), &n), "refine");
        return n;
    }
    int compress(const std::vector<int32_t> &ids)
    {
        int n = 0;
        check(ctx_.get(), gq_compress_blocks(ctx_.get(), ids.data(), (int)ids.size(), &n), "compress");
        return n;
    }
    std::vector<double> dumpUniform() const
    {
        const size_t g = (size_t)bpd_ * 16;
        std::vector<double> v(g * g * g, 0.0);
        check(ctx_.get(), gq_dump_uniform(ctx_.get(), v.data(), v.size()), "gq_dump_uniform");
        return v;
    }
    int blocksPerDimension() const { return bpd_; }
    int maxLevel() const { return maxLevel_; }
    gq_ctx *ctx() const { return ctx_.get(); }

private:
    Context &ctx_;
    int bpd_, maxLevel_;
    BlockCollection<B> coll_;
    BoundaryInfo binfo_;
};

// BlockProcessing_Pipeline_TBB drop-in (MRAGmultithreading/
// MRAGBlockProcessing_TBB.h:249-264, 476-512): pipeline_process stages the
// RD functor for a collection, the static process(UpdateTumor) executes the
// fused step on the GPU.  Staging is per host thread (the reference's
// pipeline is driven by one thread too).
class BlockProcessing {
    struct Staged {
        gq_ctx *ctx = nullptr;
        double Dw = 0, Dg = 0, rho = 0;
    };
    static Staged &staged()
    {
        static thread_local Staged s;
        return s;
    }

public:
    template <typename B, int DIM>
    void pipeline_process(const std::vector<BlockInfo> &, const BlockCollection<B> &collection, BoundaryInfo &,
                          const Glioma_ReactionDiffusionOperator<DIM> &rhs)
    {
        Staged &s = staged();
        s.ctx = collection.ctx();
        s.Dw = rhs.Dw;
        s.Dg = rhs.Dg;
        s.rho = rhs.rho;
    }
    template <typename B, int DIM>
    static void process(const std::vector<BlockInfo> &, const BlockCollection<B> &collection,
                        const UpdateTumor<DIM> &up, int /*granularity*/ = 4)
    {
        Staged &s = staged();
        gq_ctx *c = collection.ctx();
        if (!s.ctx || s.ctx != c) throw Error(GQ_ERR_STATE, "UpdateTumor without a staged RD operator");
        gq_fk_params p = {s.Dw, s.Dg, s.rho, up.dt};
        check(c, gq_fk_steps(c, &p, 1), "gq_fk_steps");
        s.ctx = nullptr;
    }
    // n fused steps in one call (the hot loop of Glioma_HG_UQ::run)
    template <typename W, typename B, int DIM>
    static void steps(Grid<W, B> &grid, const Glioma_ReactionDiffusionOperator<DIM> &rhs, const UpdateTumor<DIM> &up,
                      int n)
    {
        gq_fk_params p = {rhs.Dw, rhs.Dg, rhs.rho, up.dt};
        check(grid.ctx(), gq_fk_steps(grid.ctx(), &p, n), "gq_fk_steps");
    }
};

// Science::AutomaticRefinement / AutomaticCompression
// (MRAGscience/MRAGScienceCore.h:42-108, :117-177), with the reference's
// template arguments (channels of the projector: 0..0) and arguments; the
// refilling callback is not supported (the split transfers the data).
struct Science {
    template <int iFirstChannel, int iLastChannel, typename GridT, typename FWT>
    static int AutomaticRefinement(GridT &g, FWT &, double tol, int maxLevel = -1, int maxLoops = -1,
                                   Profiler * = nullptr, void (*fillGrid)(GridT &) = nullptr)
    {
        static_assert(iFirstChannel == 0 && iLastChannel == 0, "one projected channel (phi)");
        if (fillGrid) throw Error(GQ_ERR_UNSUPPORTED, "refinement refill callbacks are not supported");
        int n = 0;
        check(g.ctx(), gq_auto_refine(g.ctx(), tol, maxLevel, maxLoops, &n), "gq_auto_refine");
        return n;
    }
    template <int iFirstChannel, int iLastChannel, typename FWT, typename GridT>
    static int AutomaticCompression(GridT &g, FWT &, double tol, int maxLoops = -1, Profiler * = nullptr,
                                    void (*fillGrid)(GridT &) = nullptr)
    {
        static_assert(iFirstChannel == 0 && iLastChannel == 0, "one projected channel (phi)");
        if (fillGrid) throw Error(GQ_ERR_UNSUPPORTED, "compression refill callbacks are not supported");
        int n = 0;
        check(g.ctx(), gq_auto_compress(g.ctx(), tol, maxLoops, &n), "gq_auto_compress");
        return n;
    }
    // the same without the FWT argument
    template <typename GridT>
    static int AutomaticRefinement(GridT &g, double tol, int maxLevel = -1, int maxLoops = -1)
    {
        int n = 0;
        check(g.ctx(), gq_auto_refine(g.ctx(), tol, maxLevel, maxLoops, &n), "gq_auto_refine");
        return n;
    }
    template <typename GridT>
    static int AutomaticCompression(GridT &g, double tol, int maxLoops = -1)
    {
        int n = 0;
        check(g.ctx(), gq_auto_compress(g.ctx(), tol, maxLoops, &n), "gq_auto_compress");
        return n;
    }
};

}  // namespace mrag
}  // namespace gq
