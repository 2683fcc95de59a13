// gpr/api.hpp -- header-only C++ mirror of Graptor's operator API over the
// C ABI in gpr.h.
//
// Graptor (include/graptor/api.h) builds expression ASTs from lambdas:
//
//   api::edgemap( GA,
//       api::record( output, api::reduction, api::strong ),
//       api::filter( api::weak, api::src, F ),
//       api::filter( api::weak, api::dst, [&]( auto d ) { return level[d] > _c(iter+1); } ),
//       api::relax( [&]( auto s, auto d, auto e ) { return level[d].min( level[s] + _1 ); } )
//   ).materialize();
//
// Lambdas cannot cross a C ABI, so each lambda of the four drop-in
// applications is named here as a fixed operator with the same meaning:
//
//   api::edgemap( GA,
//       api::record( output, api::reduction, api::strong ),
//       api::filter( api::weak, api::src, F ),
//       api::filter( api::weak, api::dst, api::gt( level, iter+1 ) ),
//       api::relax( api::min_plus1( level ) )
//   ).materialize();
//
// Names and argument order follow api.h: filter (:687-706), record
// (:1450-1516), relax (:1670-1687), config / frac_threshold /
// always_sparse / always_dense (:130-222, :1595-1612), make_lazy_executor
// with vertex_map / vertex_scan / vertex_filter (dsl/vertexmap.h:2098-2456),
// frontier (frontier/frontier.h:58-1179), vertexprop (api/vertexprop.h).
// Errors throw gpr::error carrying gpr_last_error's text.
#pragma once

#include "../gpr.h"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace gpr {

struct error : std::runtime_error {
    int code;
    error(int c, const std::string &w) : std::runtime_error(w), code(c) {}
};

inline void check(int rc, const char *what) {
    if (rc != GPR_OK) {
        char buf[1024];
        gpr_last_error(buf, sizeof(buf));
        throw error(rc, std::string(what) + ": " + buf);
    }
}

// A graph handle (the GraphType of the reference benchmarks).
class Graph {
public:
    explicit Graph(gpr_graph *g) : g_(g) {}
    gpr_graph *get() const { return g_; }
    uint32_t numVertices() const {
        uint32_t n;
        check(gpr_graph_info(g_, &n, nullptr, nullptr), "graph info");
        return n;
    }
    uint64_t numEdges() const {
        uint64_t m;
        check(gpr_graph_info(g_, nullptr, &m, nullptr), "graph info");
        return m;
    }
    bool isSymmetric() const {
        int s;
        check(gpr_graph_info(g_, nullptr, nullptr, &s), "graph info");
        return s != 0;
    }

private:
    gpr_graph *g_;
};

namespace api {

enum class strength { weak, strong };
constexpr strength weak = strength::weak;
constexpr strength strong = strength::strong;
struct src_t {};
struct dst_t {};
constexpr src_t src{};
constexpr dst_t dst{};

// frontier: no destructor, shallow copies share storage, del() releases
// (frontier.h:266-296, 1170-1178).
class frontier {
public:
    frontier() = default;
    static frontier empty(const Graph &G) {
        frontier f;
        check(gpr_frontier_create(G.get(), &f.f_), "frontier create");
        return f;
    }
    // frontier::create(n, v, deg)
    static frontier create(const Graph &G, uint32_t v) {
        frontier f = empty(G);
        check(gpr_frontier_set_single(f.f_, v), "frontier create");
        return f;
    }
    static frontier all_true(const Graph &G) {
        frontier f = empty(G);
        check(gpr_frontier_set_all(f.f_), "frontier all_true");
        return f;
    }
    static frontier sparse(const Graph &G, const std::vector<uint32_t> &vs) {
        frontier f = empty(G);
        check(gpr_frontier_set_list(f.f_, vs.data(), (uint32_t)vs.size()), "frontier sparse");
        return f;
    }
    gpr_frontier *get() const { return f_; }
    void ensure(const Graph &G) {
        if (!f_) check(gpr_frontier_create(G.get(), &f_), "frontier create");
    }
    uint32_t nActiveVertices() const {
        uint32_t v = 0;
        if (f_) check(gpr_frontier_stats(f_, &v, nullptr, nullptr), "frontier stats");
        return v;
    }
    uint64_t nActiveEdges() const {
        uint64_t e = 0;
        if (f_) check(gpr_frontier_stats(f_, nullptr, &e, nullptr), "frontier stats");
        return e;
    }
    int getType() const {
        int k = GPR_FRONTIER_EMPTY;
        if (f_) check(gpr_frontier_stats(f_, nullptr, nullptr, &k), "frontier stats");
        return k;
    }
    bool isEmpty() const { return nActiveVertices() == 0; }
    // density (frontier.h:707-716)
    double density(uint64_t m) const { return double(nActiveVertices() + nActiveEdges()) / double(m); }
    void toSparse() { check(gpr_frontier_convert(f_, GPR_FRONTIER_LIST), "frontier toSparse"); }
    void del() {
        if (f_) gpr_frontier_destroy(f_);
        f_ = nullptr;
    }

private:
    gpr_frontier *f_ = nullptr;
};

template <class T> struct dtype_of;
template <> struct dtype_of<uint32_t> { static constexpr int v = GPR_U32; };
template <> struct dtype_of<float> { static constexpr int v = GPR_F32; };
template <> struct dtype_of<double> { static constexpr int v = GPR_F64; };
template <> struct dtype_of<uint16_t> { static constexpr int v = GPR_U16; };

// vertexprop<T>: device-resident per-vertex array, released by del().
template <class T> class vertexprop {
public:
    vertexprop() = default;
    vertexprop(const Graph &G, const char * /*name*/) {
        check(gpr_prop_create(G.get(), dtype_of<T>::v, &p_), "vertexprop");
    }
    gpr_prop *get() const { return p_; }
    void del() {
        if (p_) gpr_prop_destroy(p_);
        p_ = nullptr;
    }
    void download(T *out) const { check(gpr_prop_download(p_, out), "vertexprop download"); }
    void upload(const T *in) { check(gpr_prop_upload(p_, in), "vertexprop upload"); }

private:
    gpr_prop *p_ = nullptr;
};

// ---- operator arguments --------------------------------------------------
struct filter_src_arg {
    strength s;
    frontier *F;
};
struct dst_pred {
    int kind;
    uint32_t arg;
};
// dst predicates: p[d] > k (BFSLVLv.C:193-196), p[d] != k (CCv.C:559-564)
inline dst_pred gt(const vertexprop<uint32_t> &, uint32_t k) { return {GPR_DST_GT, k}; }
inline dst_pred ne(const vertexprop<uint32_t> &, uint32_t k) { return {GPR_DST_NE, k}; }
struct filter_dst_arg {
    strength s;
    dst_pred p;
};
inline filter_src_arg filter(strength s, src_t, frontier &F) { return {s, &F}; }
inline filter_dst_arg filter(strength s, dst_t, dst_pred p) { return {s, p}; }

struct reduction_t {};
constexpr reduction_t reduction{};
struct reduction_or_method_t {};
constexpr reduction_or_method_t reduction_or_method{};
// The record method [&](auto d){ return p[d] != prev[d]; } (BFSLVLv.C:169-172,
// CCv.C:538-540), the one method form the engine evaluates on the device.
struct method_ne {
    gpr_prop *p;
    gpr_prop *prev;
};
inline method_ne differs(vertexprop<uint32_t> &p, vertexprop<uint32_t> &prev) { return {p.get(), prev.get()}; }
struct record_arg {
    frontier *out;
    int kind;
    strength s;
    gpr_prop *prev;
};
inline record_arg record(frontier &out, reduction_t, strength s = strong) {
    return {&out, GPR_RECORD_REDUCTION, s, nullptr};
}
// api::record(output, method, strength) (api.h:1480-1496)
inline record_arg record(frontier &out, method_ne m, strength s = strong) {
    return {&out, GPR_RECORD_METHOD_NE, s, m.prev};
}
// api::record(output, api::reduction_or_method, method, strength)
// (api.h:1509-1516): the method in push, the reduction in pull
inline record_arg record(frontier &out, reduction_or_method_t, method_ne m, strength s = strong) {
    return {&out, GPR_RECORD_REDUCTION_OR_METHOD, s, m.prev};
}

struct relax_arg {
    int kind;
    gpr_prop *prop;
    gpr_prop *contrib;
    int fmt;
    float damping;
};
// level[d].min( level[s] + _1 )
inline relax_arg min_plus1(vertexprop<uint32_t> &p) { return {GPR_RELAX_MIN_PLUS1, p.get(), nullptr, 0, 0.f}; }
// IDs[d].min( IDs[s] )
inline relax_arg min_label(vertexprop<uint32_t> &p) { return {GPR_RELAX_MIN, p.get(), nullptr, 0, 0.f}; }
// parent[d].setif(s) (BFSv.C:146)
inline relax_arg setif(vertexprop<uint32_t> &p) { return {GPR_RELAX_SETIF, p.get(), nullptr, 0, 0.f}; }
// y_acc[d] += cast<acc>( contrib[s] )
template <class CT>
inline relax_arg pr_sum(vertexprop<float> &y, vertexprop<CT> &contrib, int fmt, float damping = 0.85f) {
    return {GPR_RELAX_PR_SUM, y.get(), contrib.get(), fmt, damping};
}
inline relax_arg relax(relax_arg r) { return r; }

struct frac_threshold {
    float pct;
    explicit frac_threshold(float t) : pct(t) {}
};
struct always_sparse_t {};
struct always_dense_t {};
struct learned_policy_t {};
constexpr always_sparse_t always_sparse{};
constexpr always_dense_t always_dense{};
constexpr learned_policy_t learned_policy{};
// pct 0: default_threshold (api.h:150-153); > 0: frac_threshold(pct)
struct config_arg {
    int dir = GPR_DIR_AUTO;
    float pct = 0.f;
    int out_kind = 0;
};
inline config_arg config(frac_threshold t) {
    config_arg c;
    c.pct = t.pct;
    return c;
}
inline config_arg config(always_sparse_t) {
    config_arg c;
    c.dir = GPR_DIR_PUSH;
    return c;
}
inline config_arg config(always_dense_t) {
    config_arg c;
    c.dir = GPR_DIR_PULL;
    return c;
}
inline config_arg config(learned_policy_t) {
    config_arg c;
    c.dir = GPR_DIR_POLICY;
    return c;
}
inline config_arg config(int dir, float pct, int out_kind) {
    config_arg c;
    c.dir = dir;
    c.pct = pct;
    c.out_kind = out_kind;
    return c;
}

// api::fusion / api::fusion_select (api/fusion.h:104-117, api.h:227-248): a
// sparse step that processes the vertices it activates at once.  The
// reference's fusion lambda decides per destination; for the min relaxes it
// is "process if changed", which the engine applies, so here the argument
// only selects when a step fuses.
struct fusion_arg {
    int mode;
};
inline fusion_arg fusion() { return {GPR_FUSION_DEFAULT}; }   // default_fusion_select
inline fusion_arg fusion_select(bool enabled) { return {enabled ? GPR_FUSION_ON : GPR_FUSION_OFF}; }
inline fusion_arg fusion_mode(int mode) { return {mode}; }

// ---- lazy executor (vertex ops run at materialize) ---------------------------
struct vop {
    int op;
    gpr_prop *p;
    const gpr_prop *q;
    frontier *F;
    uint64_t v;
};

class lazy_executor {
public:
    explicit lazy_executor(const Graph &G, gpr_iter_stats st = {}) : G_(G), st_(st) {}
    lazy_executor &vertex_map_fill(vertexprop<uint32_t> &p, uint32_t v) {
        ops_.push_back({GPR_VMAP_FILL, p.get(), nullptr, nullptr, v});
        return *this;
    }
    lazy_executor &vertex_map_fill(vertexprop<float> &p, float v) {
        uint32_t bits;
        static_assert(sizeof(bits) == sizeof(v), "f32");
        __builtin_memcpy(&bits, &v, 4);
        ops_.push_back({GPR_VMAP_FILL, p.get(), nullptr, nullptr, bits});
        return *this;
    }
    lazy_executor &vertex_map_iota(vertexprop<uint32_t> &p) {
        ops_.push_back({GPR_VMAP_IOTA, p.get(), nullptr, nullptr, 0});
        return *this;
    }
    lazy_executor &vertex_map_fill(frontier &F, vertexprop<uint32_t> &p, uint32_t v) {
        ops_.push_back({GPR_VMAP_FILL_FRONTIER, p.get(), nullptr, &F, v});
        return *this;
    }
    template <class T> lazy_executor &vertex_map_copy(vertexprop<T> &dst, const vertexprop<T> &src) {
        ops_.push_back({GPR_VMAP_COPY, dst.get(), src.get(), nullptr, 0});
        return *this;
    }
    // vertex_map(F, [&](auto v){ return dst[v] = src[v]; })
    lazy_executor &vertex_map_copy(frontier &F, vertexprop<uint32_t> &dst, const vertexprop<uint32_t> &src) {
        ops_.push_back({GPR_VMAP_COPY_FRONTIER, dst.get(), src.get(), &F, 0});
        return *this;
    }
    // vertex_filter(G, F, out, p[v] == value) / != value
    lazy_executor &vertex_filter(frontier &F, frontier &out, vertexprop<uint32_t> &p, uint32_t value, bool neg) {
        filters_.push_back({&F, &out, p.get(), value, neg});
        return *this;
    }
    const gpr_iter_stats &stats() const { return st_; }
    gpr_iter_stats materialize() {
        for (auto &o : ops_)
            check(gpr_vertex_map(G_.get(), o.op, o.p, o.q, o.F ? o.F->get() : nullptr, o.v), "vertex_map");
        for (auto &f : filters_) {
            f.out->ensure(G_);
            check(gpr_vertex_filter(G_.get(), f.in->get(), f.p, f.value, f.neg ? 1 : 0, f.out->get()),
                  "vertex_filter");
        }
        ops_.clear();
        filters_.clear();
        return st_;
    }

private:
    struct vfilter {
        frontier *in, *out;
        gpr_prop *p;
        uint32_t value;
        bool neg;
    };
    Graph G_;
    gpr_iter_stats st_;
    std::vector<vop> ops_;
    std::vector<vfilter> filters_;
};

inline lazy_executor make_lazy_executor(const Graph &G) { return lazy_executor(G); }

// maintain_copies (dsl/vertexmap.h:2515-2535): dst[v] = src[v] over the
// change set, or over every vertex when it holds more than n/4 of them or
// is not a vertex set.
inline void maintain_copies(const Graph &G, frontier &change, vertexprop<uint32_t> &dst,
                            const vertexprop<uint32_t> &src) {
    const int kind = change.get() ? change.getType() : GPR_FRONTIER_ALL;
    const bool do_all = kind == GPR_FRONTIER_ALL || change.nActiveVertices() > G.numVertices() / 4;
    if (do_all)
        make_lazy_executor(G).vertex_map_copy(dst, src).materialize();
    else if (!change.isEmpty())
        make_lazy_executor(G).vertex_map_copy(change, dst, src).materialize();
}

// ---- edgemap ------------------------------------------------------------------
namespace detail {
inline void apply(gpr_emap_desc &d, frontier *&in, frontier *&out, const filter_src_arg &a) {
    d.src_strong = a.s == strong;
    in = a.F;
}
inline void apply(gpr_emap_desc &d, frontier *&, frontier *&, const filter_dst_arg &a) {
    d.dst_filter = a.p.kind;
    d.dst_arg = a.p.arg;
}
inline void apply(gpr_emap_desc &d, frontier *&, frontier *&out, const record_arg &a) {
    d.record = a.kind;
    d.prev = a.prev;
    out = a.out;
}
inline void apply(gpr_emap_desc &d, frontier *&, frontier *&, const relax_arg &a) {
    d.relax = a.kind;
    d.prop = a.prop;
    d.contrib = a.contrib;
    d.contrib_fmt = a.fmt;
    d.damping = a.damping;
}
inline void apply(gpr_emap_desc &d, frontier *&, frontier *&, const config_arg &a) {
    d.direction = a.dir;
    d.threshold_pct = a.pct;
    d.out_kind = a.out_kind;
}
inline void apply(gpr_emap_desc &d, frontier *&, frontier *&, const fusion_arg &a) { d.fusion = a.mode; }
} // namespace detail

// api::edgemap: one step, executed at once; the returned executor carries
// the step statistics and runs chained vertex maps at materialize().
template <class... Args> lazy_executor edgemap(const Graph &G, Args &&...args) {
    gpr_emap_desc d = {};
    d.relax = -1;
    d.threshold_pct = 0.f;   // default_threshold unless api::config says otherwise
    frontier *in = nullptr, *out = nullptr;
    (detail::apply(d, in, out, args), ...);
    if (d.relax < 0) throw error(GPR_E_INVAL, "edgemap: must specify relax operation");
    if (out) out->ensure(G);
    gpr_iter_stats st = {};
    check(gpr_edgemap(G.get(), &d, in ? in->get() : nullptr, out ? out->get() : nullptr, &st), "edgemap");
    return lazy_executor(G, st);
}

} // namespace api
} // namespace gpr
