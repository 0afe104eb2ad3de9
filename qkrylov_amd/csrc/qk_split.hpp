// qk_split.hpp -- the two-pass apply over a relabelled spin basis.
//
// The single-pass apply walks rows in the public basis order. For a model
// such as the J1-J2 square lattice most of a row's targets then lie far away
// in that order, and each gather misses the caches. Here the sites are
// relabelled into three groups, the state read from its top bits as
// [B | C | A] (A the low NL bits, ranked per lane), chosen so that every
// off-diagonal term acts inside A u B or inside A u C:
//   pass 1: the terms inside A u B plus every diagonal term, walking the
//           tiles with the C bits fixed (all B patterns for one C pattern);
//   pass 2: the remaining terms, inside A u C, walking the tiles in memory
//           order, i.e. with the B bits (the top bits) fixed.
// Within one pass every target of a row shares the row's fixed group, so
// the rows a pass visits for one fixed pattern (a "unit", up to C(21, 10)
// rows = 2.8 MB of a vector) stay in one XCD's L2 while that XCD walks
// them. The relabelled basis is the same spin-1/2 sector; vectors in its
// row order ("internal order") are a permutation of the public ones, and
// the Krylov drivers keep their vectors in it, converting the start and the
// Ritz vector only.
#pragma once
#include "qk_engine.hpp"

#include <memory>
#include <vector>

namespace qk {

struct SplitPlan {
    std::vector<int> perm;   // public site i -> internal bit perm[i]
    int NL = 0, nB = 0, nC = 0;
    OpSum ops1, ops2;        // relabelled terms of pass 1 and pass 2
};

// Finds a split of a spin-1/2 model with nsites <= 32: A = two windows of
// w sites in site-index order (cyclic), B and C the runs between them, each
// union A u B and A u C at most max_unit sites. Returns false when no such
// split separates every off-diagonal term.
bool find_split(const BasisSpec& basis, const OpSum& ops, int max_unit, SplitPlan& out);

class SplitEngines {
public:
    SplitEngines(const Engine& pub, SplitPlan plan);
    const SplitPlan& plan() const { return plan_; }
    const Engine& pass1() const { return *e1_; }
    const Engine& pass2() const { return *e2_; }
    const QkTables* tables1() const { return &T1_; }   // pass-1 tables with the C-major tile order
    const QkTables* tables2() const { return e2_->tables(); }
    // y = H x in internal order (tmp: one vector of scratch)
    void apply(const double* x, double* y, double* tmp, hipStream_t s) const;
    // launch tuning of both passes: knob 0 tile schedule, 2 waves per
    // workgroup (as Engine::set_tuning)
    void set_tuning(int knob, int value) const;
    int tuning(int knob) const { return e2_->tuning(knob); }
    // public order <-> internal order (one vector each)
    void to_internal(const double* x_pub, double* x_int, hipStream_t s) const;
    void to_public(const double* x_int, double* x_pub, hipStream_t s) const;

private:
    const Engine& pub_;
    SplitPlan plan_;
    std::shared_ptr<const BasisSpec> basis_;
    std::unique_ptr<Engine> e1_, e2_;
    DevBuf tiles1_, bitmap_;
    mutable QkTables T1_{};
};

} // namespace qk
