// mgpcg_batched.hip -- the batched device MG-PCG behind the feat3_batch_* C ABI: K systems
// that share one sparsity graph solved in lockstep (same algorithm as mgpcg.hip, per-system
// scalars and statuses; reference: solver/pcg.hpp:188-288, multigrid.hpp:1268-2000,
// richardson.hpp:187-228, jacobi_precond.hpp:137-165, convert_precond.hpp:100-119).
//
//  * K right-hand sides of one matrix (values shared), e.g. 32 RHS of the 3x3-block vector
//    Laplacian (SparseMatrixBCSR<3,3>, sparse_matrix_bcsr.hpp:1295-1600; Jacobi on the
//    diagonal of the diagonal blocks, diagonal_generic.hpp:39-58);
//  * K coefficient fields (one value set per system), e.g. 16 variable-coefficient diffusion
//    matrices on one mesh.
// Layout in HBM: graph (u32 row_ptr/col_ind) once; values [entry][h][w] or [entry][h][w][k];
// vectors [node][h][k]; transfers as scalar graphs applied per component (compact, no stored
// zeros); node Dirichlet bitmask; per-system PCG scalars (SysState[K]).

#include "batch_kernels.hpp"
#include "device_common.hpp"
#include "mg_kernels.hpp"
#include "runtime.hpp"
#include "feat3_hip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

using namespace feat3_amd;
using feat3_amd::dev::half;
typedef std::uint32_t u32;
typedef std::uint64_t u64;

namespace
{
  constexpr int kBlock = 256;
#define F3_CHECK(...) do { if((__VA_ARGS__) != hipSuccess) return rt::kErrRuntime; } while(0)
#define F3_RC(...) do { const int rc_ = (__VA_ARGS__); if(rc_) return rc_; } while(0)

  struct DevMem
  {
    void* p = nullptr;
    ~DevMem() { if(p) (void)hipFree(p); }
    template<typename T> T* as() const { return static_cast<T*>(p); }
    int alloc(u64 bytes)
    {
      if(p) { (void)hipFree(p); p = nullptr; }
      return hipMalloc(&p, bytes ? bytes : 16) == hipSuccess ? rt::kOk : rt::kErrRuntime;
    }
    template<typename T> int put(const std::vector<T>& h, hipStream_t s)
    {
      F3_RC(alloc(sizeof(T) * h.size()));
      if(!h.empty()) F3_CHECK(hipMemcpyAsync(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
      return rt::kOk;
    }
  };

  template<typename VT> VT to_storage(double v) { return VT(v); }
  template<> half to_storage<half>(double v) { return half(float(v)); }

  struct HostGraph
  {
    u64 rows = 0, cols = 0;
    std::vector<u32> rp, col;
    std::vector<double> val;   // nnz * vals_per_entry
  };

  struct HostLevel
  {
    u64 nodes = 0;
    HostGraph A, P, R;
    std::vector<u32> fidx;
  };

  struct Config
  {
    double omega_j = 0.5, omega_r = 1.0;
    int pre = 4, post = 4;
    double tol_rel = 1e-8, tol_abs = 1.0 / (2.220446049250313e-16 * 2.220446049250313e-16), tol_abs_low = 0.0;
    u64 min_iter = 0, max_iter = 1000;
    bool use_graph = true;
    u64 tile_rows = 0;     // rows per locality tile of the row kernels (0/1: ascending order)
    // shared B x B values whose every block is diagonal are stored as block diagonals (the
    // off-diagonal exact zeros left out, as the transfers' structural zeros are)
    bool diag_blocks = true;
    // shared values: the levels of at most this many nodes (below the finest) run in one
    // LDS-resident workgroup per system (batch::k_tail_sys); 0 = every level keeps its launches
    u64 tail_nodes = 1000;
  };

  // drop stored exact zeros of a scalar graph (the transfers' structural zeros)
  HostGraph compact(const HostGraph& h)
  {
    HostGraph c;
    c.rows = h.rows; c.cols = h.cols;
    c.rp.assign(h.rows + 1, 0);
    for(u64 r = 0; r < h.rows; ++r)
    {
      for(u32 k = h.rp[r]; k < h.rp[r + 1]; ++k)
        if(h.val[k] != 0.0) { c.col.push_back(h.col[k]); c.val.push_back(h.val[k]); }
      c.rp[r + 1] = u32(c.col.size());
    }
    return c;
  }

  // The inner-precision graph of shared block values (FP32 / binary16 arithmetic) leaves out the
  // entries whose largest value is at most eps_f32^2 of the largest diagonal value of their row's
  // diagonal block: such a block changes a row sum by less than eps^2 of the diagonal term, far
  // below the rounding of FP32 arithmetic (the face couplings of the 3D Q1 vector Laplacian are
  // roundoff residues of about 6e-17 of the diagonal: 21 of 27 blocks per row stay).  The outer
  // FP64 product keeps the full graph.  A block holding a non-finite value is kept, and so is
  // every block of a row with a non-finite diagonal value: the inner copy of a non-finite matrix
  // stays non-finite.
  HostGraph thin_graph(const HostGraph& h, u64 bb, int B)
  {
    const double eps2 = 1.1920928955078125e-07 * 1.1920928955078125e-07;
    HostGraph c;
    c.rows = h.rows; c.cols = h.cols;
    c.rp.assign(h.rows + 1, 0);
    for(u64 r = 0; r < h.rows; ++r)
    {
      double d = 0.0;
      bool finite_diag = true;
      for(u32 k = h.rp[r]; k < h.rp[r + 1]; ++k)
        if(u64(h.col[k]) == r)
          for(int q = 0; q < B; ++q)
          {
            const double v = h.val[u64(k) * bb + u64(q) * u64(B) + u64(q)];
            finite_diag = finite_diag && std::isfinite(v);
            d = std::max(d, std::fabs(v));
          }
      for(u32 k = h.rp[r]; k < h.rp[r + 1]; ++k)
      {
        double m = 0.0;
        bool finite = finite_diag;
        for(u64 j = 0; j < bb; ++j)
        {
          const double v = h.val[u64(k) * bb + j];
          finite = finite && std::isfinite(v);
          m = std::max(m, std::fabs(v));
        }
        if(u64(h.col[k]) != r && finite && m <= eps2 * d) continue;
        c.col.push_back(h.col[k]);
        for(u64 j = 0; j < bb; ++j) c.val.push_back(h.val[u64(k) * bb + j]);
      }
      c.rp[r + 1] = u32(c.col.size());
    }
    return c;
  }

  bool invert_dense(std::vector<double>& a, u64 n)
  {
    std::vector<double> inv(n * n, 0.0);
    for(u64 i = 0; i < n; ++i) inv[i * n + i] = 1.0;
    for(u64 k = 0; k < n; ++k)
    {
      u64 pr = k;
      for(u64 i = k + 1; i < n; ++i) if(std::fabs(a[i * n + k]) > std::fabs(a[pr * n + k])) pr = i;
      if(a[pr * n + k] == 0.0) return false;
      if(pr != k) for(u64 j = 0; j < n; ++j) { std::swap(a[k * n + j], a[pr * n + j]); std::swap(inv[k * n + j], inv[pr * n + j]); }
      const double d = a[k * n + k];
      for(u64 j = 0; j < n; ++j) { a[k * n + j] /= d; inv[k * n + j] /= d; }
      for(u64 i = 0; i < n; ++i)
      {
        if(i == k) continue;
        const double f = a[i * n + k];
        if(f == 0.0) continue;
        for(u64 j = 0; j < n; ++j) { a[i * n + j] -= f * a[k * n + j]; inv[i * n + j] -= f * inv[k * n + j]; }
      }
    }
    a.swap(inv);
    return true;
  }

  inline unsigned stream_blocks(u64 items, int G)
  {
    const u64 per_block = 256 / u64(G);
    return unsigned(std::max<u64>(1, std::min<u64>((items + per_block - 1) / per_block, 4096)));
  }

  struct BatchBase
  {
    virtual ~BatchBase() {}
    virtual int init(const std::vector<HostLevel>& lv, const Config& c, hipStream_t s) = 0;
    virtual int solve(double* x, const double* b, int* status, u64* iters) = 0;
    virtual int history(int k, double* out, u64 cap, u64* len) = 0;
    virtual int time_kernel(int which, int reps, double* ms, u64* bytes) = 0;
    virtual bool set_tolerances(const Config& c) = 0;
    virtual int set_values(int level, const double* dvals) = 0;
  };

  // device-side refresh of a level's values (feat3_batch_set_values): storage conversion and
  // the Jacobi diagonal with the host setup's arithmetic (diagonal rounded to the storage type,
  // dinv = omega_R * (omega_J / diag) in the inner type, 0 on Dirichlet nodes)
  template<typename VT>
  __global__ void k_to_storage(u64 n, const double* __restrict__ src, VT* __restrict__ dst)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      dst[i] = VT(float(src[i]));
  }
  template<>
  __global__ void k_to_storage<double>(u64 n, const double* __restrict__ src, double* __restrict__ dst)
  {
    for(u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x)
      dst[i] = src[i];
  }

  template<typename DT, typename VT, int B>
  __global__ void k_batch_dinv(u64 nodes, u64 kv, const u32* __restrict__ rp, const u32* __restrict__ col,
                               const double* __restrict__ val, const u32* __restrict__ mask, DT omega_j, DT omega_r,
                               DT* __restrict__ dinv)
  {
    const u64 total = nodes * u64(B) * kv;
    for(u64 t = u64(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += u64(gridDim.x) * blockDim.x)
    {
      const u64 k = t % kv;
      const u64 nh = t / kv;
      const u64 node = nh / u64(B);
      const u64 hh = nh % u64(B);
      DT d = DT(0);
      for(u32 e = rp[node]; e < rp[node + 1]; ++e)
        if(u64(col[e]) == node)
        {
          const double v = val[(u64(e) * u64(B * B) + hh * u64(B) + hh) * kv + k];
          if constexpr(sizeof(VT) == 8) d = DT(v);
          else d = DT(VT(float(v)));
          break;
        }
      dinv[t] = dev::masked(mask, node) ? DT(0) : omega_r * (omega_j / d);
    }
  }

  template<int N>
  using ic = std::integral_constant<int, N>;

  // Lane layouts.  Streaming kernels: one system per lane, G = 16/32/64 lanes per node
  // component.  Row kernels over vectors of type T: S = 16 / sizeof(T) systems per lane (one
  // 16-byte access) and GV = K / S lanes per row when K is a multiple of S and K / S a power of
  // two; otherwise S = 1 and GV = G.
  inline int stream_lanes(int K) { return K <= 16 ? 16 : (K <= 32 ? 32 : 64); }

  template<typename T>
  void row_layout(int K, int& GV, int& S)
  {
    const int sv = int(16 / sizeof(T));
    const int q = K / sv;
    if(K % sv == 0 && q >= 1 && (q & (q - 1)) == 0) { GV = q; S = sv; }
    else { GV = stream_lanes(K); S = 1; }
  }

  template<typename T, typename F>
  int with_layout(int GV, int S, F&& f)
  {
    constexpr int SV = int(16 / sizeof(T));
    bool done = false;
    auto one = [&](auto gv, auto sv) {
      if(!done && GV == decltype(gv)::value && S == decltype(sv)::value) { f(gv, sv); done = true; }
    };
    one(ic<16>(), ic<1>()); one(ic<32>(), ic<1>()); one(ic<64>(), ic<1>());
    one(ic<1>(), ic<SV>()); one(ic<2>(), ic<SV>()); one(ic<4>(), ic<SV>()); one(ic<8>(), ic<SV>()); one(ic<16>(), ic<SV>());
    if constexpr(SV == 2) one(ic<32>(), ic<SV>());
    return done ? rt::kOk : rt::kErrArg;
  }

  // grid of the V-cycle's row kernels: one pass of 256 / G rows per block, so that a block pays
  // its rows' dependent round trips (row pointer, columns, gathers) once instead of once per
  // 64-row pass (mg::row_grid), which left the small levels latency-bound.  Only the V-cycle's
  // kernels use it; they reduce nothing (no per-block partials), so the results do not depend
  // on it.
  inline mg::RowGrid vcycle_grid(u64 rows, int G)
  {
    const u64 per = 256 / u64(G);
    const u64 nb = (rows + per - 1) / per;
    if(nb >= (u64(1) << 31)) return mg::row_grid(rows, G);
    mg::RowGrid g;
    g.per = per;
    g.blocks = unsigned(nb < 1 ? 1 : nb);
    return g;
  }

  template<typename F>
  int with_stream(int G, F&& f)
  {
    if(G == 16) f(ic<16>());
    else if(G == 32) f(ic<32>());
    else if(G == 64) f(ic<64>());
    else return rt::kErrArg;
    return rt::kOk;
  }

  // DIAG (shared 3x3 values whose every block is diagonal, e.g. a vector Laplacian): the device
  // copies of the level matrices and of the FP64 outer matrix hold only the block diagonals
  template<typename DT, typename VT, int B, bool SHARED, bool DIAG = false>
  struct Engine : public BatchBase
  {
    struct Level
    {
      u64 nodes = 0, nnz = 0;
      DevMem arp, acol, aval, prp, pcol, pval, rrp, rcol, rval, mask, dinv, rhs, x, t, def, ord;
      u64 nc = 0;
      bool thinned = false;   // the graph holds fewer entries than the pushed one (thin_graph)
      // ELL copies of A, P, R for the coarse tail (levels tail0 and below), widths aw, pw, rw
      DevMem ea_col, ea_val, ep_col, ep_val, er_col, er_val;
      u32 aw = 0, pw = 0, rw = 0;
      DT* xfinal = nullptr;
      const u32* order() const { return ord.p ? ord.template as<u32>() : nullptr; }
    };
    Config cfg;
    hipStream_t s = nullptr;
    int K = 1;
    int G = 16, GV = 16, S = 1, GO = 16, SO = 1;   // streaming, V-cycle rows, outer FP64 rows
    std::vector<std::unique_ptr<Level>> L;
    DevMem ainv, a64, r, p, q, st, cr, hist, partials, sums, any;
    DevMem orp, ocol;   // the outer product's graph where level 0's inner graph is thinned
    u64 m0 = 0, hist_cap = 0, ncoarse_rows = 0;
    // coarse tail: levels tail0..nl-1 run in k_tail_sys when tail0 + 1 < nl (shared values);
    // tail_lds bytes of LDS hold one system's vectors on them
    std::size_t tail0 = 0;
    u64 tail_lds = 0;
    // published system states (batch::k_publish_states): a ring of kRing slots of 64 states in
    // pinned, host-coherent memory (ring_dev: its device address), pub_count the device's count
    static constexpr int kRing = 4;
    struct HostRing
    {
      batch::SysState st[kRing][64];
      unsigned seq[kRing];
    };
    HostRing* ring = nullptr;
    HostRing* ring_dev = nullptr;
    DevMem pub_count;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    // graphs: [0] the first outer half (q = A p, alpha, x and r, criteria) and the publication;
    // [1] the preconditioner half of iteration k (V-cycle, beta, p), the outer half of k + 1 and
    // the publication
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    void drop_graphs()
    {
      for(int i = 0; i < 2; ++i)
      {
        if(gexec[i]) { (void)hipGraphExecDestroy(gexec[i]); gexec[i] = nullptr; }
        if(graph[i]) { (void)hipGraphDestroy(graph[i]); graph[i] = nullptr; }
      }
    }
    double* x_cur = nullptr;
    double* x_graph = nullptr;
    std::vector<u64> last_iters;
    HostLevel coarse_host;   // graph and Dirichlet list of the coarsest level (coarse inverses)

    explicit Engine(int k) : K(k)
    {
      G = stream_lanes(K);
      row_layout<DT>(K, GV, S);
      row_layout<double>(K, GO, SO);
    }
    ~Engine() override
    {
      drop_graphs();
      if(ring) (void)hipHostFree(ring);
      if(t0) (void)hipEventDestroy(t0);
      if(t1) (void)hipEventDestroy(t1);
    }

    static constexpr u64 vpe() { return u64(B) * u64(B); }   // values per entry and system (as pushed)
    static constexpr u64 svpe() { return DIAG ? u64(B) : vpe(); }   // as stored on the device
    u64 vals_per_entry() const { return svpe() * (SHARED ? 1 : u64(K)); }
    // the block diagonals of full B x B blocks (DIAG); other layouts unchanged
    static std::vector<double> stored(const std::vector<double>& v)
    {
      if(!DIAG) return v;
      std::vector<double> o(v.size() / vpe() * B);
      for(std::size_t e = 0; e < v.size() / vpe(); ++e)
        for(int h = 0; h < B; ++h) o[e * B + h] = v[e * vpe() + u64(h) * B + u64(h)];
      return o;
    }

    template<typename T> std::vector<T> conv(const std::vector<double>& v) const
    {
      std::vector<T> o(v.size());
      for(std::size_t i = 0; i < v.size(); ++i) o[i] = to_storage<T>(v[i]);
      return o;
    }

    template<int S_> batch::BlockRows<DT, VT, B, SHARED, S_, DIAG> rowsA(const Level& l) const
    {
      return {l.arp.template as<u32>(), l.acol.template as<u32>(), l.aval.template as<VT>(), l.nodes, K, l.order()};
    }
    // P: rows are the fine nodes of level l; R: rows are the coarse nodes (level c)
    template<int S_> batch::CompRows<DT, VT, B, S_> rowsP(const Level& l) const
    {
      return {l.prp.template as<u32>(), l.pcol.template as<u32>(), l.pval.template as<VT>(), l.nodes, K, l.order()};
    }
    template<int S_> batch::CompRows<DT, VT, B, S_> rowsR(const Level& l, const Level& c) const
    {
      return {l.rrp.template as<u32>(), l.rcol.template as<u32>(), l.rval.template as<VT>(), l.nc, K, c.order()};
    }
    template<int S_> batch::BlockRows<double, double, B, SHARED, S_, DIAG> rows64() const
    {
      const bool own = L[0]->thinned;
      return {(own ? orp : L[0]->arp).template as<u32>(), (own ? ocol : L[0]->acol).template as<u32>(), a64.template as<double>(),
              L[0]->nodes, K,
              L[0]->order()};
    }

    int init(const std::vector<HostLevel>& lv, const Config& c, hipStream_t stream) override
    {
      cfg = c;
      s = stream;
      if(lv.empty() || cfg.pre < 1 || cfg.post < 1) return rt::kErrArg;
      L.clear();
      for(std::size_t i = 0; i < lv.size(); ++i)
      {
        const HostLevel& h = lv[i];
        std::unique_ptr<Level> l(new Level());
        l->nodes = h.nodes;
        if(SHARED && sizeof(VT) < 8)
        {
          const HostGraph T = thin_graph(h.A, vpe(), B);
          l->thinned = T.rp[h.nodes] != h.A.rp[h.nodes];
          l->nnz = T.rp[h.nodes];
          F3_RC(l->arp.put(T.rp, s));
          F3_RC(l->acol.put(T.col, s));
          F3_RC(l->aval.put(conv<VT>(stored(T.val)), s));
          F3_CHECK(hipStreamSynchronize(s));   // T leaves scope
        }
        else
        {
          l->nnz = h.A.rp[h.nodes];
          F3_RC(l->arp.put(h.A.rp, s));
          F3_RC(l->acol.put(h.A.col, s));
          F3_RC(l->aval.put(conv<VT>(stored(h.A.val)), s));
        }
        if(i + 1 < lv.size())
        {
          const HostGraph P = compact(h.P), R = compact(h.R);
          l->nc = lv[i + 1].nodes;
          F3_RC(l->prp.put(P.rp, s)); F3_RC(l->pcol.put(P.col, s)); F3_RC(l->pval.put(conv<VT>(P.val), s));
          F3_RC(l->rrp.put(R.rp, s)); F3_RC(l->rcol.put(R.col, s)); F3_RC(l->rval.put(conv<VT>(R.val), s));
        }
        std::vector<u32> m((h.nodes + 31) / 32 + 1, 0u);
        for(u32 f : h.fidx) m[f >> 5] |= 1u << (f & 31u);
        F3_RC(l->mask.put(m, s));
        if(cfg.tile_rows > 1 && h.nodes > cfg.tile_rows)
        {
          std::vector<u32> ord(h.nodes);
          if(feat3_setup_tile_order(h.nodes, h.A.rp.data(), h.A.col.data(), cfg.tile_rows, ord.data())) return rt::kErrArg;
          F3_RC(l->ord.put(ord, s));
        }
        // Jacobi: dinv = omega_R * omega_J / diag (diagonal of the diagonal block), 0 at Dirichlet
        const u64 kv = SHARED ? 1 : u64(K);
        std::vector<DT> dinv(h.nodes * B * kv, DT(0));
        for(u64 node = 0; node < h.nodes; ++node)
        {
          const bool dir = (m[node >> 5] >> (node & 31u)) & 1u;
          for(u32 e = h.A.rp[node]; e < h.A.rp[node + 1]; ++e)
            if(h.A.col[e] == node)
              for(int hh = 0; hh < B; ++hh)
                for(u64 k = 0; k < kv; ++k)
                {
                  // the diagonal in the storage precision, as the reference's Jacobi extracts it
                  const DT d = DT(to_storage<VT>(h.A.val[(u64(e) * vpe() + u64(hh) * B + u64(hh)) * kv + k]));
                  dinv[(node * B + u64(hh)) * kv + k] = dir ? DT(0) : DT(cfg.omega_r) * (DT(cfg.omega_j) / d);
                }
        }
        F3_RC(l->dinv.put(dinv, s));
        const u64 nv = h.nodes * B * u64(K);
        F3_RC(l->rhs.alloc(sizeof(DT) * nv)); F3_RC(l->x.alloc(sizeof(DT) * nv));
        F3_RC(l->t.alloc(sizeof(DT) * nv)); F3_RC(l->def.alloc(sizeof(DT) * nv));
        L.push_back(std::move(l));
      }
      // FP64 inverses of the unit-filtered coarse matrices
      coarse_host = lv.back();
      F3_RC(coarse_inverses(coarse_host.A.val));
      for(std::size_t i = 0; i + 1 < L.size(); ++i)
        L[i]->xfinal = ((cfg.pre - 1 + cfg.post) % 2 == 0) ? L[i]->x.template as<DT>() : L[i]->t.template as<DT>();
      L.back()->xfinal = L.back()->x.template as<DT>();
      F3_RC(setup_tail());
      return init_outer(lv[0]);
    }

    // the deepest run of levels (the coarsest included) whose every level has at most
    // cfg.tail_nodes nodes and whose vectors fit one workgroup's LDS
    int setup_tail()
    {
      const std::size_t nl = L.size();
      tail0 = nl - 1;
      tail_lds = 0;
      if(!SHARED || cfg.tail_nodes == 0) return rt::kOk;
      int dev = 0, cap = 65536;
      if(hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cap, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
      auto bytes = [&](std::size_t t0) {
        u64 b = 0;
        for(std::size_t i = t0; i < nl; ++i) b += 5 * L[i]->nodes * u64(B) * sizeof(DT);
        return b;
      };
      while(tail0 > 0 && L[tail0 - 1]->nodes <= cfg.tail_nodes && nl - (tail0 - 1) <= std::size_t(batch::kTailSysMax) &&
            bytes(tail0 - 1) <= u64(cap))
        --tail0;
      if(tail0 + 1 >= nl) return rt::kOk;
      tail_lds = bytes(tail0);
      for(std::size_t i = tail0; i < nl; ++i)
      {
        Level& l = *L[i];
        F3_RC(tail_ell(l.arp, l.acol, l.aval, l.nodes, u32(svpe()), l.ea_col, l.ea_val, l.aw));
        if(i + 1 < nl)
        {
          F3_RC(tail_ell(l.prp, l.pcol, l.pval, l.nodes, 1u, l.ep_col, l.ep_val, l.pw));
          F3_RC(tail_ell(l.rrp, l.rcol, l.rval, l.nc, 1u, l.er_col, l.er_val, l.rw));
        }
      }
      F3_CHECK(hipStreamSynchronize(s));
      if constexpr(SHARED)
      {
        if(tail_lds > 65536)
          F3_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&batch::k_tail_sys<DT, VT, B, DIAG>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(tail_lds)));
      }
      return rt::kOk;
    }

    // ELL copy (batch::TailSysEll) of a device CSR matrix of `rows` rows with V values per
    // entry; w = its longest row.  Rebuilt from the CSR values when they change (set_values).
    int tail_ell(const DevMem& rp, const DevMem& col, const DevMem& val, u64 rows, u32 V, DevMem& ecol, DevMem& eval, u32& w)
    {
      std::vector<u32> h(rows + 1);
      F3_CHECK(hipMemcpyAsync(h.data(), rp.p, sizeof(u32) * (rows + 1), hipMemcpyDeviceToHost, s));
      F3_CHECK(hipStreamSynchronize(s));
      w = 0;
      for(u64 r = 0; r < rows; ++r) w = std::max(w, h[r + 1] - h[r]);
      if(w == 0) w = 1;
      if(!ecol.p) F3_RC(ecol.alloc(sizeof(u32) * rows * w));
      if(!eval.p) F3_RC(eval.alloc(sizeof(VT) * rows * w * V));
      const u64 n = rows * w;
      const unsigned nb = unsigned(std::max<u64>(1, std::min<u64>((n + 255) / 256, 65536)));
      batch::k_tail_ell<VT><<<nb, 256, 0, s>>>(u32(rows), w, V, rp.template as<u32>(), col.template as<u32>(),
                                               val.template as<VT>(), ecol.template as<u32>(), eval.template as<VT>());
      return rt::check_launch();
    }

    // the tail's levels as the kernel reads them (current device pointers: set_values may
    // replace the values and the coarse inverse)
    batch::TailSysDesc<DT, VT> tail_desc() const
    {
      batch::TailSysDesc<DT, VT> d{};
      const std::size_t nl = L.size();
      d.nlev = int(nl - tail0);
      d.pre = cfg.pre;
      d.post = cfg.post;
      d.ncoarse = u32(ncoarse_rows);
      d.ainv = ainv.template as<double>();
      d.rhs_in = L[tail0]->rhs.template as<DT>();
      d.x_out = L[tail0]->xfinal;
      u64 off = 0;
      for(std::size_t i = tail0; i < nl; ++i)
      {
        const Level& l = *L[i];
        batch::TailSysLevel<DT, VT>& t = d.lv[i - tail0];
        t.nodes = u32(l.nodes);
        t.off = u32(off);
        off += 5 * l.nodes * u64(B);
        t.A = {l.arp.template as<u32>(), l.ea_col.template as<u32>(), l.ea_val.template as<VT>(), u32(l.nodes), l.aw};
        if(i + 1 < nl)
        {
          t.P = {l.prp.template as<u32>(), l.ep_col.template as<u32>(), l.ep_val.template as<VT>(), u32(l.nodes), l.pw};
          t.R = {l.rrp.template as<u32>(), l.er_col.template as<u32>(), l.er_val.template as<VT>(), u32(l.nc), l.rw};
        }
        t.mask = l.mask.template as<u32>();
        t.dinv = l.dinv.template as<DT>();
      }
      return d;
    }

    // FP64 inverses of the unit-filtered coarse matrices from the coarsest level's values
    int coarse_inverses(const std::vector<double>& cval)
    {
      const HostLevel& hc = coarse_host;
      ncoarse_rows = hc.nodes * B;
      const u64 mc = ncoarse_rows;
      if(mc > 4096) return rt::kErrArg;
      const u64 ninv = SHARED ? 1 : u64(K);
      std::vector<double> inv_all(ninv * mc * mc);
      for(u64 k = 0; k < ninv; ++k)
      {
        std::vector<double> a(mc * mc, 0.0);
        for(u64 node = 0; node < hc.nodes; ++node)
          for(u32 e = hc.A.rp[node]; e < hc.A.rp[node + 1]; ++e)
            for(int hh = 0; hh < B; ++hh)
              for(int ww = 0; ww < B; ++ww)
                a[(node * B + hh) * mc + u64(hc.A.col[e]) * B + ww] +=
                  cval[(u64(e) * vpe() + u64(hh) * B + u64(ww)) * ninv + k];
        for(u32 f : hc.fidx)
          for(int hh = 0; hh < B; ++hh)
          {
            const u64 row = u64(f) * B + hh;
            for(u64 j = 0; j < mc; ++j) a[row * mc + j] = 0.0;
            a[row * mc + row] = 1.0;
          }
        if(!invert_dense(a, mc)) return rt::kErrNumeric;
        std::copy(a.begin(), a.end(), inv_all.begin() + k * mc * mc);
      }
      F3_RC(ainv.put(inv_all, s));
      F3_CHECK(hipStreamSynchronize(s));   // inv_all goes out of scope
      return rt::kOk;
    }

    int init_outer(const HostLevel& l0)
    {
      // outer FP64
      m0 = l0.nodes * B;
      F3_RC(a64.put(stored(l0.A.val), s));
      if(L[0]->thinned)
      {
        F3_RC(orp.put(l0.A.rp, s));
        F3_RC(ocol.put(l0.A.col, s));
      }
      const u64 nv = m0 * u64(K);
      F3_RC(r.alloc(sizeof(double) * nv)); F3_RC(p.alloc(sizeof(double) * nv)); F3_RC(q.alloc(sizeof(double) * nv));
      F3_RC(st.alloc(sizeof(batch::SysState) * K));
      F3_RC(cr.alloc(sizeof(batch::Criteria)));
      hist_cap = cfg.max_iter + 2;
      F3_RC(hist.alloc(sizeof(double) * hist_cap * K));
      F3_RC(partials.alloc(sizeof(double) * 8192 * u64(K)));
      F3_RC(sums.alloc(sizeof(double) * u64(K)));
      F3_RC(any.alloc(64));
      if(!ring)
      {
        F3_CHECK(hipHostMalloc(reinterpret_cast<void**>(&ring), sizeof(HostRing), hipHostMallocMapped | hipHostMallocCoherent));
        std::memset(ring, 0, sizeof(HostRing));
        F3_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&ring_dev), ring, 0));
      }
      F3_RC(pub_count.alloc(sizeof(unsigned)));
      if(!t0) F3_CHECK(hipEventCreate(&t0));
      if(!t1) F3_CHECK(hipEventCreate(&t1));
      F3_CHECK(hipStreamSynchronize(s));
      drop_graphs();
      return rt::kOk;
    }

    bool set_tolerances(const Config& c) override
    {
      if(c.max_iter + 2 > hist_cap) return false;
      cfg.tol_rel = c.tol_rel; cfg.tol_abs = c.tol_abs; cfg.tol_abs_low = c.tol_abs_low;
      cfg.min_iter = c.min_iter; cfg.max_iter = c.max_iter;
      return true;
    }

    const int* anyp() const { return any.template as<int>(); }

    int sweep(Level& f, DT* xin, DT* xout, const int* a)
    {
      return with_layout<DT>(GV, S, [&](auto gv, auto sv) {
        constexpr int Gc = decltype(gv)::value, Sc = decltype(sv)::value;
        const mg::RowGrid g = vcycle_grid(f.nodes, Gc);
        batch::k_sweep<Gc, Sc, B, SHARED><<<g.blocks, kBlock, 0, s>>>(rowsA<Sc>(f), g.per, f.dinv.template as<DT>(),
                                                                      f.rhs.template as<DT>(), (const DT*)xin, xout, a);
      });
    }

    int defect(Level& f, const DT* x, const int* a)
    {
      return with_layout<DT>(GV, S, [&](auto gv, auto sv) {
        constexpr int Gc = decltype(gv)::value, Sc = decltype(sv)::value;
        const mg::RowGrid g = vcycle_grid(f.nodes, Gc);
        batch::k_defect<Gc, Sc, B><<<g.blocks, kBlock, 0, s>>>(rowsA<Sc>(f), g.per, f.mask.template as<u32>(),
                                                              f.rhs.template as<DT>(), x, f.def.template as<DT>(), a);
      });
    }

    int restrict_to(Level& f, Level& c, DT* xc, const int* a)
    {
      return with_layout<DT>(GV, S, [&](auto gv, auto sv) {
        constexpr int Gc = decltype(gv)::value, Sc = decltype(sv)::value;
        const mg::RowGrid g = vcycle_grid(f.nc, Gc);
        batch::k_restrict<Gc, Sc, B, SHARED><<<g.blocks, kBlock, 0, s>>>(rowsR<Sc>(f, c), g.per, c.mask.template as<u32>(),
                                                                         f.def.template as<DT>(), c.rhs.template as<DT>(), xc,
                                                                         c.dinv.template as<DT>(), a);
      });
    }

    int prolong(Level& f, const DT* xc, DT* xf, const int* a)
    {
      return with_layout<DT>(GV, S, [&](auto gv, auto sv) {
        constexpr int Gc = decltype(gv)::value, Sc = decltype(sv)::value;
        const mg::RowGrid g = vcycle_grid(f.nodes, Gc);
        batch::k_prolong<Gc, Sc, B><<<g.blocks, kBlock, 0, s>>>(rowsP<Sc>(f), g.per, f.mask.template as<u32>(), xc, xf, a);
      });
    }

    // entered: rhs and x of the top level already hold the cycle's entry (written by k_pcg_update)
    int enqueue_vcycle(const double* def, const int* a, bool entered = false)
    {
      const std::size_t nl = L.size();
      Level& top = *L[0];
      if(!entered)
        batch::k_entry<DT, B, SHARED><<<4096, kBlock, 0, s>>>(top.nodes, K, def, top.rhs.template as<DT>(), top.x.template as<DT>(),
                                                             top.dinv.template as<DT>(), a);
      std::vector<DT*> cur(nl, nullptr), oth(nl, nullptr);
      // levels above the coarse tail (every level but the coarsest when there is none); the
      // tail takes its first level's first smoothing step itself
      for(std::size_t i = 0; i < tail0; ++i)
      {
        Level& f = *L[i];
        Level& c = *L[i + 1];
        cur[i] = f.x.template as<DT>();
        oth[i] = f.t.template as<DT>();
        for(int k = 1; k < cfg.pre; ++k) { F3_RC(sweep(f, cur[i], oth[i], a)); std::swap(cur[i], oth[i]); }
        F3_RC(defect(f, cur[i], a));
        F3_RC(restrict_to(f, c, (i + 1 < tail0) ? c.x.template as<DT>() : nullptr, a));
      }
      if constexpr(SHARED)
      {
        if(tail0 + 1 < nl)
          batch::k_tail_sys<DT, VT, B, DIAG><<<unsigned(K), batch::kTailSysThreads, tail_lds, s>>>(tail_desc(), K, a);
      }
      if(tail0 + 1 >= nl)
      {
        Level& cl = *L.back();
        const unsigned cb = unsigned(std::max<u64>(1, (ncoarse_rows * u64(K) + kBlock - 1) / kBlock));
        batch::k_coarse<DT><<<cb, kBlock, 0, s>>>(ncoarse_rows, K, SHARED ? 0 : 1, ainv.template as<double>(),
                                                 cl.mask.template as<u32>(), B, cl.rhs.template as<DT>(), cl.x.template as<DT>(), a);
      }
      for(std::size_t ii = tail0; ii > 0; --ii)
      {
        const std::size_t i = ii - 1;
        Level& f = *L[i];
        Level& c = *L[i + 1];
        F3_RC(prolong(f, c.xfinal, cur[i], a));
        for(int k = 0; k < cfg.post; ++k) { F3_RC(sweep(f, cur[i], oth[i], a)); std::swap(cur[i], oth[i]); }
        if(cur[i] != f.xfinal) return rt::kErrState;
      }
      return rt::check_launch();
    }

    // finalisation of a per-system reduction whose producing kernel ran with nb blocks: the K
    // sums (one block per system), then the K systems' scalars (one block)
    template<int OP>
    void fin(unsigned nb)
    {
      batch::k_fin_sum<<<unsigned(K), kBlock, 0, s>>>(partials.template as<double>(), nb, K, sums.template as<double>());
      batch::k_fin<OP><<<1, kBlock, 0, s>>>(sums.template as<double>(), K, st.template as<batch::SysState>(),
                                             cr.template as<batch::Criteria>(), hist.template as<double>(), any.template as<int>());
    }

    // outer q = filter(A p) with partials of q.p; returns the number of blocks (the partial count)
    int outer_spmv_dot(unsigned& nb)
    {
      return with_layout<double>(GO, SO, [&](auto gv, auto sv) {
        constexpr int Gc = decltype(gv)::value, Sc = decltype(sv)::value;
        const mg::RowGrid g = mg::row_grid(L[0]->nodes, Gc);
        nb = g.blocks;
        batch::k_pcg_spmv_dot<Gc, Sc, B><<<g.blocks, kBlock, 0, s>>>(rows64<Sc>(), g.per, L[0]->mask.template as<u32>(),
                                                                    p.template as<double>(), q.template as<double>(),
                                                                    st.template as<batch::SysState>(), anyp(),
                                                                    partials.template as<double>());
      });
    }

    int outer_defect(const double* b, const double* x, unsigned& nb)
    {
      return with_layout<double>(GO, SO, [&](auto gv, auto sv) {
        constexpr int Gc = decltype(gv)::value, Sc = decltype(sv)::value;
        const mg::RowGrid g = mg::row_grid(L[0]->nodes, Gc);
        nb = g.blocks;
        batch::k_pcg_defect<Gc, Sc, B><<<g.blocks, kBlock, 0, s>>>(rows64<Sc>(), g.per, L[0]->mask.template as<u32>(), b, x,
                                                                  r.template as<double>(), partials.template as<double>());
      });
    }

    // outer half of a PCG iteration: q = A p, alpha, x and r updates, the criteria
    int enqueue_outer_half()
    {
      double* part = partials.template as<double>();
      batch::SysState* Sst = st.template as<batch::SysState>();
      unsigned nb = 0;
      F3_RC(outer_spmv_dot(nb));
      fin<mg::kFinAlpha>(nb);
      const unsigned sb = stream_blocks(m0, G);
      F3_RC(with_stream(G, [&](auto gg) {
        constexpr int Gc = decltype(gg)::value;
        batch::k_pcg_update<Gc, B, SHARED, DT><<<sb, kBlock, 0, s>>>(L[0]->nodes, K, p.template as<double>(), q.template as<double>(),
                                                                     x_cur, r.template as<double>(), Sst, anyp(), part,
                                                                     L[0]->rhs.template as<DT>(), L[0]->x.template as<DT>(),
                                                                     L[0]->dinv.template as<DT>());
      }));
      fin<mg::kFinNewDefect>(sb);
      return rt::check_launch();
    }

    // preconditioner half: z = MG(r), beta, p = z + beta p
    int enqueue_precond_half()
    {
      double* part = partials.template as<double>();
      batch::SysState* Sst = st.template as<batch::SysState>();
      const unsigned sb = stream_blocks(m0, G);
      const DT* z = L[0]->xfinal;
      // the outer half before it (k_pcg_update) wrote the cycle's entry
      F3_RC(enqueue_vcycle(r.template as<double>(), anyp(), true));
      F3_RC(with_stream(G, [&](auto gg) {
        constexpr int Gc = decltype(gg)::value;
        batch::k_pcg_rz<Gc, B, DT><<<sb, kBlock, 0, s>>>(L[0]->nodes, K, z, r.template as<double>(), Sst, anyp(), part);
      }));
      fin<mg::kFinBeta>(sb);
      F3_RC(with_stream(G, [&](auto gg) {
        constexpr int Gc = decltype(gg)::value;
        batch::k_pcg_update_p<Gc, B, DT><<<sb, kBlock, 0, s>>>(L[0]->nodes, K, z, p.template as<double>(), Sst, anyp());
      }));
      return rt::check_launch();
    }

    int solve(double* x, const double* b, int* status, u64* iters) override
    {
      if(!x || !b) return rt::kErrArg;
      // the row kernels access S consecutive systems as one 16-byte vector
      if((reinterpret_cast<std::uintptr_t>(x) | reinterpret_cast<std::uintptr_t>(b)) % 16 != 0) return rt::kErrArg;
      batch::Criteria c{};
      c.tol_rel = cfg.tol_rel; c.tol_abs = cfg.tol_abs; c.tol_abs_low = cfg.tol_abs_low;
      c.div_rel = 1.0 / 2.220446049250313e-16; c.div_abs = 1.0 / (2.220446049250313e-16 * 2.220446049250313e-16);
      c.stag_rate = 0.95; c.min_iter = cfg.min_iter; c.max_iter = cfg.max_iter; c.min_stag_iter = 0; c.hist_cap = hist_cap;
      F3_CHECK(hipMemcpyAsync(cr.p, &c, sizeof(c), hipMemcpyHostToDevice, s));
      x_cur = x;
      double* part = partials.template as<double>();
      batch::SysState* Sst = st.template as<batch::SysState>();
      unsigned nb = 0;
      F3_RC(outer_defect(b, x, nb));
      fin<mg::kFinInitDefect>(nb);
      F3_RC(enqueue_vcycle(r.template as<double>(), anyp()));
      const unsigned sb = stream_blocks(m0, G);
      F3_RC(with_stream(G, [&](auto gg) {
        constexpr int Gc = decltype(gg)::value;
        batch::k_pcg_p_gamma<Gc, B, DT><<<sb, kBlock, 0, s>>>(L[0]->nodes, K, L[0]->xfinal, r.template as<double>(),
                                                              p.template as<double>(), Sst, part);
      }));
      fin<mg::kFinGamma>(sb);
      F3_RC(rt::check_launch());
      F3_CHECK(hipMemsetAsync(pub_count.p, 0, sizeof(unsigned), s));
      F3_RC(enqueue_publish());   // publication 1: the initial defects
      auto enqueue_graph = [&](int g) -> int {
        if(g == 0) { F3_RC(enqueue_outer_half()); return enqueue_publish(); }
        F3_RC(enqueue_precond_half());
        F3_RC(enqueue_outer_half());
        return enqueue_publish();
      };
      if(cfg.use_graph && (gexec[0] == nullptr || x_graph != x))
      {
        drop_graphs();
        for(int g = 0; g < 2; ++g)
        {
          F3_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
          const int rc = enqueue_graph(g);
          hipGraph_t gr = nullptr;
          const hipError_t e = hipStreamEndCapture(s, &gr);
          if(rc) { if(gr) (void)hipGraphDestroy(gr); return rc; }
          if(e != hipSuccess) return rt::kErrRuntime;
          graph[g] = gr;
          F3_CHECK(hipGraphInstantiate(&gexec[g], graph[g], nullptr, nullptr, 0));
        }
        x_graph = x;
      }
      unsigned launched = 1, read = 0;   // publications launched and read (in order)
      auto launch = [&](int g) -> int {
        if(cfg.use_graph) F3_CHECK(hipGraphLaunch(gexec[g], s));
        else F3_RC(enqueue_graph(g));
        ++launched;
        return rt::kOk;
      };
      // defects of every system after iterations 0..read-1 (-1: stopped), as read by the host
      std::vector<std::vector<double>> seen;
      std::vector<batch::SysState> last(K);
      bool running = true;
      auto read_state = [&]() -> int {
        const unsigned want = read + 1, slot = (want - 1) % kRing;
        for(unsigned spin = 1; __atomic_load_n(&ring->seq[slot], __ATOMIC_ACQUIRE) != want; ++spin)
          if(spin % 256 == 0)
          {
            // the wait ends: once the stream has drained, the publication has arrived or is lost
            const hipError_t qs = hipStreamQuery(s);
            if(qs == hipSuccess)
            {
              if(__atomic_load_n(&ring->seq[slot], __ATOMIC_ACQUIRE) == want) break;
              return rt::kErrRuntime;
            }
            if(qs != hipErrorNotReady) return rt::kErrRuntime;
          }
        std::vector<double> d(K);
        running = false;
        for(int j = 0; j < K; ++j)
        {
          last[j] = ring->st[slot][j];
          d[j] = last[j].status == mg::kProgress ? last[j].def_cur : -1.0;
          running = running || last[j].status == mg::kProgress;
        }
        seen.push_back(std::move(d));
        read = want;
        return rt::kOk;
      };
      // Host loop, as the single-system solver's: the next iteration is launched while the
      // previous one runs; when every running system's defect history predicts that the outer
      // half just launched may meet its criterion, its states are read before the next graph is
      // launched
      const u64 max_launch = cfg.max_iter + 1;
      u64 k = 0;          // outer halves launched
      bool wait = false;  // the latest outer half may end every running system
      while(k < max_launch && running)
      {
        if(wait)
        {
          while(read < launched && running) F3_RC(read_state());
          if(!running) break;
        }
        F3_RC(launch(k == 0 ? 0 : 1));
        ++k;
        wait = k >= cfg.max_iter;
        if(!wait && seen.size() >= 2)
        {
          const std::size_t jj = seen.size() - 1;
          wait = true;
          for(int j = 0; j < K && wait; ++j)
          {
            const double dj = seen[jj][j], dp = seen[jj - 1][j], d0 = seen[0][j];
            if(dj < 0.0) continue;   // this system has stopped
            const double rate = dp > 0.0 ? dj / dp : 1.0;
            const double goal = std::min(cfg.tol_abs, std::max(cfg.tol_rel * d0, cfg.tol_abs_low));
            wait = dj * std::pow(std::min(rate, 1.0), double(k - jj)) <= 10.0 * goal;
          }
        }
        if(!wait)
          while(read + 1 < launched && running) F3_RC(read_state());
      }
      while(read < launched) F3_RC(read_state());   // the last publication holds the final states
      F3_CHECK(hipStreamSynchronize(s));
      last_iters.assign(K, 0);
      for(int j = 0; j < K; ++j)
      {
        if(status) status[j] = last[j].status;
        if(iters) iters[j] = last[j].num_iter;
        last_iters[j] = last[j].num_iter;
      }
      return rt::kOk;
    }

    int enqueue_publish()
    {
      batch::k_publish_states<kRing><<<1, 64, 0, s>>>(st.template as<batch::SysState>(), K, &ring_dev->st[0][0], ring_dev->seq,
                                                      pub_count.template as<unsigned>());
      return rt::check_launch();
    }

    int history(int k, double* out, u64 cap, u64* len) override
    {
      if(k < 0 || k >= K || last_iters.empty()) return rt::kErrArg;
      const u64 n = std::min<u64>(last_iters[k] + 1, hist_cap);
      const u64 m = std::min(cap, n);
      if(m) F3_CHECK(hipMemcpy(out, hist.template as<double>() + u64(k) * hist_cap, sizeof(double) * m, hipMemcpyDeviceToHost));
      if(len) *len = n;
      return rt::kOk;
    }

    int set_values(int level, const double* dv) override
    {
      if(level < 0 || level >= int(L.size()) || !dv) return rt::kErrArg;
      Level& l = *L[level];
      // a thinned graph or the block-diagonal storage was chosen for the values given at init;
      // other values need a new solver
      if(l.thinned || DIAG) return rt::kErrState;
      const u64 nval = l.nnz * vals_per_entry();
      const unsigned nb = unsigned(std::max<u64>(1, std::min<u64>((nval + 255) / 256, 65536)));
      k_to_storage<VT><<<nb, 256, 0, s>>>(nval, dv, l.aval.template as<VT>());
      const u64 kv = SHARED ? 1 : u64(K);
      const u64 nd = l.nodes * B * kv;
      const unsigned nbd = unsigned(std::max<u64>(1, std::min<u64>((nd + 255) / 256, 65536)));
      k_batch_dinv<DT, VT, B><<<nbd, 256, 0, s>>>(l.nodes, kv, l.arp.template as<u32>(), l.acol.template as<u32>(), dv,
                                                   l.mask.template as<u32>(), DT(cfg.omega_j), DT(cfg.omega_r),
                                                   l.dinv.template as<DT>());
      F3_RC(rt::check_launch());
      if(tail0 + 1 < L.size() && std::size_t(level) >= tail0)
        F3_RC(tail_ell(l.arp, l.acol, l.aval, l.nodes, u32(svpe()), l.ea_col, l.ea_val, l.aw));
      if(level == 0) F3_CHECK(hipMemcpyAsync(a64.p, dv, sizeof(double) * nval, hipMemcpyDeviceToDevice, s));
      if(level + 1 == int(L.size()))
      {
        std::vector<double> cval(nval);
        F3_CHECK(hipMemcpyAsync(cval.data(), dv, sizeof(double) * nval, hipMemcpyDeviceToHost, s));
        F3_CHECK(hipStreamSynchronize(s));
        F3_RC(coarse_inverses(cval));
      }
      F3_CHECK(hipStreamSynchronize(s));
      return rt::kOk;
    }

    // which: 0 finest smoothing sweep (all systems), 2 outer FP64 q = A p with dot, 3 one V-cycle
    int time_kernel(int which, int reps, double* ms, u64* bytes) override
    {
      if(reps < 1 || (which != 0 && which != 2 && which != 3)) return rt::kErrArg;
      Level& f = *L[0];
      if(which == 2)
      {
        // a state where every system is in progress, so that the kernel does its full work
        std::vector<batch::SysState> sp(K);
        for(auto& v : sp) { v = batch::SysState{}; v.status = mg::kProgress; }
        const int one = 1;
        F3_CHECK(hipStreamSynchronize(s));
        F3_CHECK(hipMemcpy(st.p, sp.data(), sizeof(batch::SysState) * K, hipMemcpyHostToDevice));
        F3_CHECK(hipMemcpy(any.p, &one, sizeof(int), hipMemcpyHostToDevice));
      }
      for(int pass = 0; pass < 2; ++pass)
      {
        const int n = pass == 0 ? 1 : reps;
        if(pass == 1) F3_CHECK(hipEventRecord(t0, s));
        for(int k = 0; k < n; ++k)
        {
          unsigned nb = 0;
          if(which == 0) F3_RC(sweep(f, f.x.template as<DT>(), f.t.template as<DT>(), nullptr));
          else if(which == 2) F3_RC(outer_spmv_dot(nb));
          else F3_RC(enqueue_vcycle(r.template as<double>(), nullptr));
        }
        if(pass == 1) F3_CHECK(hipEventRecord(t1, s));
        F3_CHECK(hipStreamSynchronize(s));
      }
      float e = 0.f;
      F3_CHECK(hipEventElapsedTime(&e, t0, t1));
      *ms = double(e) / reps;
      if(bytes)
      {
        // algorithmic: values + column indices + row pointers, x once, and the row-wise vectors
        // (sweep: rhs, x_out, dinv; outer: p read for the dot is the x read, q written)
        const u64 nvk = f.nodes * B * u64(K);
        if(which == 2)
          *bytes = f.nnz * (vals_per_entry() * 8 + 4) + (f.nodes + 1) * 4 + nvk * 8 * 2;
        else
          *bytes = f.nnz * (vals_per_entry() * sizeof(VT) + 4) + (f.nodes + 1) * 4 + nvk * sizeof(DT) * 3 +
                   f.nodes * B * (SHARED ? 1 : u64(K)) * sizeof(DT);
      }
      return rt::kOk;
    }
  };

  template<typename DT, typename VT>
  BatchBase* make_engine(int block, int K, bool shared, bool diag)
  {
    if(block == 1 && !shared) return new Engine<DT, VT, 1, false>(K);
    if(block == 1 && shared) return new Engine<DT, VT, 1, true>(K);
    if(block == 3 && shared && diag) return new Engine<DT, VT, 3, true, true>(K);
    if(block == 3 && shared) return new Engine<DT, VT, 3, true>(K);
    return nullptr;
  }

  // every block of every level matrix (shared B x B values) has exact zeros off its diagonal
  bool all_blocks_diagonal(const std::vector<HostLevel>& lv, int B)
  {
    const u64 bb = u64(B) * u64(B);
    for(const HostLevel& l : lv)
      for(std::size_t i = 0; i < l.A.val.size(); ++i)
        if(i % bb % u64(B + 1) != 0 && l.A.val[i] != 0.0) return false;
    return true;
  }
} // namespace

struct feat3_batch
{
  int inner = 0, block = 1, K = 1;
  bool shared = true;
  Config cfg;
  std::vector<HostLevel> levels;
  std::unique_ptr<BatchBase> eng;
  hipStream_t stream = nullptr;
};

extern "C" {

int feat3_batch_create(feat3_batch** h, int inner_precision, int block, int nsys, int shared_values)
{
  if(!h || inner_precision < 0 || inner_precision > 2 || nsys < 1 || nsys > 64) return rt::kErrArg;
  if(!((block == 3 && shared_values) || block == 1)) return rt::kErrArg;
  std::unique_ptr<feat3_batch> o(new feat3_batch());
  o->inner = inner_precision; o->block = block; o->K = nsys; o->shared = shared_values != 0;
  if(hipStreamCreateWithFlags(&o->stream, hipStreamNonBlocking) != hipSuccess) return rt::kErrRuntime;
  *h = o.release();
  return rt::kOk;
}

int feat3_batch_destroy(feat3_batch* h)
{
  if(!h) return rt::kOk;
  if(h->stream) (void)hipStreamSynchronize(h->stream);
  h->eng.reset();
  if(h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return rt::kOk;
}

int feat3_batch_push_level(feat3_batch* h, u64 nodes, const u32* row_ptr, const u32* col_ind, const double* val,
                           u64 nfilt, const u32* filter_nodes, u64 nc, const u32* p_row_ptr, const u32* p_col_ind,
                           const double* p_val, const u32* r_row_ptr, const u32* r_col_ind, const double* r_val)
{
  if(!h || !row_ptr || (row_ptr[nodes] > 0 && (!col_ind || !val)) || (nfilt && !filter_nodes)) return rt::kErrArg;
  HostLevel L;
  L.nodes = nodes;
  L.A.rows = L.A.cols = nodes;
  L.A.rp.assign(row_ptr, row_ptr + nodes + 1);
  const u64 nnz = row_ptr[nodes];
  L.A.col.assign(col_ind, col_ind + nnz);
  for(u32 c : L.A.col) if(c >= nodes) return rt::kErrArg;
  const u64 vpe = u64(h->block) * u64(h->block) * (h->shared ? 1 : u64(h->K));
  L.A.val.assign(val, val + nnz * vpe);
  L.fidx.assign(filter_nodes, filter_nodes + nfilt);
  for(u32 f : L.fidx) if(f >= nodes) return rt::kErrArg;
  if(nc > 0)
  {
    if(!p_row_ptr || !p_col_ind || !p_val || !r_row_ptr || !r_col_ind || !r_val) return rt::kErrArg;
    L.P.rows = nodes; L.P.cols = nc;
    L.P.rp.assign(p_row_ptr, p_row_ptr + nodes + 1);
    L.P.col.assign(p_col_ind, p_col_ind + p_row_ptr[nodes]);
    L.P.val.assign(p_val, p_val + p_row_ptr[nodes]);
    L.R.rows = nc; L.R.cols = nodes;
    L.R.rp.assign(r_row_ptr, r_row_ptr + nc + 1);
    L.R.col.assign(r_col_ind, r_col_ind + r_row_ptr[nc]);
    L.R.val.assign(r_val, r_val + r_row_ptr[nc]);
    for(u32 c : L.P.col) if(c >= nc) return rt::kErrArg;
    for(u32 c : L.R.col) if(c >= nodes) return rt::kErrArg;
  }
  if(!h->levels.empty() && h->levels.back().P.cols != nodes) return rt::kErrArg;
  h->levels.push_back(std::move(L));
  h->eng.reset();
  return rt::kOk;
}

int feat3_batch_set_smoother(feat3_batch* h, double omega_jacobi, double omega_richardson, int pre, int post)
{
  if(!h || pre < 1 || post < 1 || !(omega_jacobi > 0.0)) return rt::kErrArg;
  h->cfg.omega_j = omega_jacobi; h->cfg.omega_r = omega_richardson; h->cfg.pre = pre; h->cfg.post = post;
  h->eng.reset();
  return rt::kOk;
}

int feat3_batch_set_tolerances(feat3_batch* h, double tol_rel, double tol_abs, double tol_abs_low, u64 min_iter, u64 max_iter)
{
  if(!h || max_iter == 0) return rt::kErrArg;
  h->cfg.tol_rel = tol_rel;
  if(tol_abs > 0) h->cfg.tol_abs = tol_abs;
  h->cfg.tol_abs_low = tol_abs_low;
  h->cfg.min_iter = min_iter;
  h->cfg.max_iter = max_iter;
  if(h->eng && !h->eng->set_tolerances(h->cfg)) h->eng.reset();
  return rt::kOk;
}

int feat3_batch_set_graph(feat3_batch* h, int graph)
{
  if(!h) return rt::kErrArg;
  h->cfg.use_graph = graph != 0;
  h->eng.reset();
  return rt::kOk;
}

int feat3_batch_set_option(feat3_batch* h, const char* name, double value)
{
  if(!h || !name) return rt::kErrArg;
  const std::string n(name);
  if(n == "tile_rows" && value >= 0) h->cfg.tile_rows = u64(value);
  else if(n == "diag_blocks") h->cfg.diag_blocks = value != 0.0;
  else if(n == "tail_max_n" && value >= 0) h->cfg.tail_nodes = u64(value);
  else return rt::kErrArg;
  h->eng.reset();
  return rt::kOk;
}

int feat3_batch_init(feat3_batch* h)
{
  if(!h || h->levels.empty()) return rt::kErrState;
  std::unique_ptr<BatchBase> e;
  const bool diag = h->cfg.diag_blocks && h->block > 1 && h->shared && all_blocks_diagonal(h->levels, h->block);
  if(h->inner == 0) e.reset(make_engine<double, double>(h->block, h->K, h->shared, diag));
  else if(h->inner == 1) e.reset(make_engine<float, float>(h->block, h->K, h->shared, diag));
  else e.reset(make_engine<float, half>(h->block, h->K, h->shared, diag));
  if(!e) return rt::kErrArg;
  const int rc = e->init(h->levels, h->cfg, h->stream);
  if(rc) return rc;
  h->eng = std::move(e);
  return rt::kOk;
}

int feat3_batch_correct(feat3_batch* h, void* caller_stream, double* x, const double* b, int* status, u64* iters)
{
  if(!h || !h->eng) return rt::kErrState;
  // the caller's stream (NULL: the null stream) has produced x and b: wait for it
  if(hipStreamSynchronize(static_cast<hipStream_t>(caller_stream)) != hipSuccess) return rt::kErrRuntime;
  return h->eng->solve(x, b, status, iters);
}

int feat3_batch_set_values(feat3_batch* h, int level, const double* values)
{
  if(!h || !h->eng) return rt::kErrState;
  // values produced on the null stream are complete before the solver's stream reads them
  if(hipStreamSynchronize(nullptr) != hipSuccess) return rt::kErrRuntime;
  return h->eng->set_values(level, values);
}

int feat3_batch_history(feat3_batch* h, int system, double* out, u64 cap, u64* len)
{
  if(!h || !h->eng) return rt::kErrState;
  return h->eng->history(system, out, cap, len);
}

int feat3_batch_time_kernel(feat3_batch* h, int which, int reps, double* avg_ms, u64* bytes)
{
  if(!h || !h->eng || !avg_ms) return rt::kErrState;
  return h->eng->time_kernel(which, reps, avg_ms, bytes);
}

} // extern "C"
