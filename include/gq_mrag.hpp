// gq_mrag.hpp -- header-only C++ mirror of the reference's block-operator API
// over the C-ABI of gq_api.h.  A driver written against MRAG + the Glioma
// operators (Glioma/Glioma_HG_UQ.cpp) keeps its shape:
//
//   Grid<W,B> grid(bpd,bpd,bpd)              -> gq::mrag::Grid grid(ctx, bpd, maxLevel, order)
//   grid.getBlocksInfo()                     -> grid.getBlocksInfo()
//   BlockProcessing::pipeline_process(vInfo, collection, boundaryInfo, rhs)
//   BlockProcessing::process(vInfo, collection, updateTumor, granularity)
//                                            -> same names; the pair runs one fused step
//   Science::AutomaticRefinement<0,0>(grid, fwt, tol, maxLevel, loops)
//   Science::AutomaticCompression<0,0>(grid, fwt, tol, loops)
//                                            -> gq::mrag::Science::Automatic{Refinement,Compression}
//   Glioma_HG_UQ::_dumpUQoutput(grid)        -> grid.dumpUniform()
//
// Errors become gq::mrag::Error exceptions carrying the gq_status.
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

// Glioma_ReactionDiffusionOperator<3>(Dw, Dg, rho): stencil [-1,2)
struct Glioma_ReactionDiffusionOperator {
    int stencil_start[3] = {-1, -1, -1};
    int stencil_end[3] = {2, 2, 2};
    double Dw, Dg, rho;
    Glioma_ReactionDiffusionOperator(double Dw_, double Dg_, double rho_) : Dw(Dw_), Dg(Dg_), rho(rho_) {}
};

// UpdateTumor<3>(dt)
struct UpdateTumor {
    double dt;
    explicit UpdateTumor(double dt_) : dt(dt_) {}
};

class Grid {
public:
    Grid(Context &ctx, int bpd, int maxLevel, int waveletOrder = 2) : ctx_(ctx), bpd_(bpd)
    {
        check(ctx.get(), gq_grid_create(ctx.get(), bpd, maxLevel, waveletOrder), "gq_grid_create");
    }
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
    gq_ctx *ctx() const { return ctx_.get(); }

private:
    Context &ctx_;
    int bpd_;
};

// BlockProcessing_Pipeline_TBB drop-in: pipeline_process stages the RD
// functor, process(UpdateTumor) executes the fused step on the GPU.
class BlockProcessing {
public:
    void pipeline_process(const std::vector<BlockInfo> &, Grid &grid, void * /*boundaryInfo*/,
                          const Glioma_ReactionDiffusionOperator &rhs)
    {
        staged_ = &grid;
        rhs_ = rhs;
    }
    void process(const std::vector<BlockInfo> &, Grid &grid, const UpdateTumor &up, int /*granularity*/ = 4)
    {
        if (staged_ != &grid) throw Error(GQ_ERR_STATE, "UpdateTumor without a staged RD operator");
        gq_fk_params p = {rhs_.Dw, rhs_.Dg, rhs_.rho, up.dt};
        check(grid.ctx(), gq_fk_steps(grid.ctx(), &p, 1), "gq_fk_steps");
        staged_ = nullptr;
    }
    // n fused steps in one call (the hot loop of Glioma_HG_UQ::run)
    static void steps(Grid &grid, const Glioma_ReactionDiffusionOperator &rhs, const UpdateTumor &up, int n)
    {
        gq_fk_params p = {rhs.Dw, rhs.Dg, rhs.rho, up.dt};
        check(grid.ctx(), gq_fk_steps(grid.ctx(), &p, n), "gq_fk_steps");
    }

private:
    Grid *staged_ = nullptr;
    Glioma_ReactionDiffusionOperator rhs_{0, 0, 0};
};

struct Science {
    static int AutomaticRefinement(Grid &g, double tol, int maxLevel = -1, int maxLoops = -1)
    {
        int n = 0;
        check(g.ctx(), gq_auto_refine(g.ctx(), tol, maxLevel, maxLoops, &n), "gq_auto_refine");
        return n;
    }
    static int AutomaticCompression(Grid &g, double tol, int maxLoops = -1)
    {
        int n = 0;
        check(g.ctx(), gq_auto_compress(g.ctx(), tol, maxLoops, &n), "gq_auto_compress");
        return n;
    }
};

}  // namespace mrag
}  // namespace gq
