// lafem_hip_driver.cpp -- C++ application-level use of the Mem::HIP containers and solvers,
// written the way applications/poisson_dirichlet.cpp:131-200 builds its solver:
//   PCG( MultiGrid-V( Richardson(1.0, Jacobi(omega)) x4 pre/post, coarse = the smoother ) )
// on 2D Q1 Poisson (levels `fine` .. `coarse`, ExpBubble load).  Runs
//   generic64  -- op-by-op solver classes, FP64 containers
//   mixed32    -- FP64 PCG + ConvertPrecond + FP32 multigrid
//   fused64    -- the fused device solver (exact coarse solve)
// and writes "<name> <status> <iterations>" lines to stdout and each solution as raw FP64 to
// <outdir>/<name>.bin for the test to compare against the oracle.
#include "feat3_amd/lafem_hip.hpp"
#include "feat3_amd/solver_hip.hpp"
#include "feat3_amd/solver_factory_hip.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

using namespace feat3_amd;
typedef std::uint32_t u32;

struct HostLevel
{
  Index n = 0, nc = 0;
  std::vector<u32> rp, col, bnd, prp, pcol, rrp, rcol;
  std::vector<double> val, pval, rval;
};

static HostLevel make_level(int level, bool with_transfer)
{
  HostLevel L;
  std::uint64_t n, nnz, nb;
  hip_check(feat3_setup_laplace_sizes(2, 1, level, &n, &nnz, &nb), "sizes");
  L.n = n;
  L.rp.resize(n + 1); L.col.resize(nnz); L.val.resize(nnz); L.bnd.resize(nb);
  hip_check(feat3_setup_laplace(2, 1, level, nullptr, L.rp.data(), L.col.data(), L.val.data(), L.bnd.data()), "laplace");
  if(with_transfer)
  {
    std::uint64_t nf, nc, pn;
    hip_check(feat3_setup_prol_sizes(2, 1, level, &nf, &nc, &pn), "prol sizes");
    L.nc = nc;
    L.prp.resize(nf + 1); L.pcol.resize(pn); L.pval.resize(pn);
    hip_check(feat3_setup_prol(2, 1, level, L.prp.data(), L.pcol.data(), L.pval.data()), "prol");
    L.rrp.resize(nc + 1); L.rcol.resize(pn); L.rval.resize(pn);
    hip_check(feat3_setup_transpose(nf, nc, L.prp.data(), L.pcol.data(), L.pval.data(), L.rrp.data(), L.rcol.data(),
                                    L.rval.data()), "transpose");
  }
  return L;
}

template<typename DT_>
struct DeviceLevel
{
  typedef LAFEM::SparseMatrixCSR<Mem::HIP, DT_, u32> Matrix;
  typedef LAFEM::UnitFilter<Mem::HIP, DT_, u32> Filter;
  typedef LAFEM::Transfer<Matrix> Transfer;
  Matrix A;
  Filter F;
  std::unique_ptr<Transfer> T;
  explicit DeviceLevel(const HostLevel& h)
    : A(h.n, h.n, h.rp, h.col, h.val), F(h.bnd)
  {
    if(h.nc)
      T.reset(new Transfer(Matrix(h.n, h.nc, h.prp, h.pcol, h.pval), Matrix(h.nc, h.n, h.rrp, h.rcol, h.rval)));
  }
};

template<typename DT_>
std::shared_ptr<Solver::MultiGrid<typename DeviceLevel<DT_>::Matrix, typename DeviceLevel<DT_>::Filter,
                                  typename DeviceLevel<DT_>::Transfer>>
make_mg(std::vector<std::unique_ptr<DeviceLevel<DT_>>>& lv, double omega,
        std::shared_ptr<Solver::MultiGridHierarchy<typename DeviceLevel<DT_>::Matrix, typename DeviceLevel<DT_>::Filter,
                                                   typename DeviceLevel<DT_>::Transfer>>& hier)
{
  typedef typename DeviceLevel<DT_>::Matrix M;
  typedef typename DeviceLevel<DT_>::Filter F;
  typedef typename DeviceLevel<DT_>::Transfer T;
  hier = std::make_shared<Solver::MultiGridHierarchy<M, F, T>>();
  for(std::size_t i = 0; i < lv.size(); ++i)
  {
    auto jac = std::make_shared<Solver::JacobiPrecond<M, F>>(lv[i]->A, lv[i]->F, DT_(omega));
    auto sm = std::make_shared<Solver::Richardson<M, F>>(lv[i]->A, lv[i]->F, DT_(1), jac);
    sm->set_min_iter(4);
    sm->set_max_iter(4);
    if(i + 1 < lv.size()) hier->push_level(lv[i]->A, lv[i]->F, *lv[i]->T, sm, sm, sm);
    else hier->push_level(lv[i]->A, lv[i]->F, sm);
  }
  return std::make_shared<Solver::MultiGrid<M, F, T>>(hier);
}

static void write(const std::string& path, const std::vector<double>& v)
{
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(double)));
}

int main(int argc, char** argv)
{
  if(argc < 5)
  {
    std::fprintf(stderr, "usage: %s <fine level> <coarse level> <jacobi omega> <outdir>\n", argv[0]);
    return 2;
  }
  const int fine = std::atoi(argv[1]), coarse = std::atoi(argv[2]);
  const double omega = std::atof(argv[3]);
  const std::string outdir = argv[4];
  try
  {
    MemoryPool<Mem::HIP>::initialise(0);
    std::vector<HostLevel> host;
    for(int l = fine; l >= coarse; --l) host.push_back(make_level(l, l > coarse));
    const HostLevel& h0 = host.front();
    std::vector<double> rhs(h0.n);
    hip_check(feat3_setup_rhs(2, 1, fine, 2, rhs.data()), "rhs");   // ExpBubble
    for(u32 i : h0.bnd) rhs[i] = 0.0;

    {
      // generic FP64
      std::vector<std::unique_ptr<DeviceLevel<double>>> lv;
      for(const HostLevel& h : host) lv.emplace_back(new DeviceLevel<double>(h));
      std::shared_ptr<Solver::MultiGridHierarchy<DeviceLevel<double>::Matrix, DeviceLevel<double>::Filter,
                                                 DeviceLevel<double>::Transfer>> hier;
      auto mg = make_mg<double>(lv, omega, hier);
      Solver::PCG<DeviceLevel<double>::Matrix, DeviceLevel<double>::Filter> pcg(lv[0]->A, lv[0]->F, mg);
      pcg.set_tol_rel(1e-8);
      pcg.set_max_iter(1000);
      hier->init();
      pcg.init();
      LAFEM::DenseVector<Mem::HIP, double, u32> x(h0.n, 0.0), b(rhs);
      lv[0]->F.filter_sol(x);
      const Solver::Status st = Solver::solve(pcg, x, b, lv[0]->A, lv[0]->F);
      std::cout << "generic64 " << int(st) << " " << pcg.get_num_iter() << std::endl;
      write(outdir + "/generic64.bin", x.to_host());

      // FP64 PCG with an FP32 multigrid through ConvertPrecond
      std::vector<std::unique_ptr<DeviceLevel<float>>> lv32;
      for(const HostLevel& h : host) lv32.emplace_back(new DeviceLevel<float>(h));
      std::shared_ptr<Solver::MultiGridHierarchy<DeviceLevel<float>::Matrix, DeviceLevel<float>::Filter,
                                                 DeviceLevel<float>::Transfer>> hier32;
      auto mg32 = make_mg<float>(lv32, omega, hier32);
      typedef LAFEM::DenseVector<Mem::HIP, double, u32> V64;
      typedef LAFEM::DenseVector<Mem::HIP, float, u32> V32;
      auto conv = std::make_shared<Solver::ConvertPrecond<V64, V32>>(mg32);
      Solver::PCG<DeviceLevel<double>::Matrix, DeviceLevel<double>::Filter> pcg2(lv[0]->A, lv[0]->F, conv);
      pcg2.set_tol_rel(1e-8);
      pcg2.set_max_iter(1000);
      hier32->init();
      pcg2.init();
      V64 x2(h0.n, 0.0);
      const Solver::Status st2 = Solver::solve(pcg2, x2, b, lv[0]->A, lv[0]->F);
      std::cout << "mixed32 " << int(st2) << " " << pcg2.get_num_iter() << std::endl;
      write(outdir + "/mixed32.bin", x2.to_host());
    }
    {
      // the fused device solver (exact FP64 coarse solve on the coarsest level)
      Solver::FusedMGPCG fused(Solver::FusedMGPCG::fp64, Solver::FusedMGPCG::sell);
      for(const HostLevel& h : host)
        fused.push_level(h.n, h.rp, h.col, h.val, h.bnd, h.nc, h.prp, h.pcol, h.pval, h.rrp, h.rcol, h.rval);
      fused.set_smoother(omega, 1.0, 4, 4);
      fused.set_tol_rel(1e-8);
      fused.set_max_iter(1000);
      fused.init();
      LAFEM::DenseVector<Mem::HIP, double, u32> x(h0.n, 0.0), b(rhs);
      const Solver::Status st = fused.correct(x, b);
      std::cout << "fused64 " << int(st) << " " << fused.get_num_iter() << std::endl;
      write(outdir + "/fused64.bin", x.to_host());
    }
    {
      // the same fused solver built by the SolverFactory from an INI configuration
      const std::string ini = "[linsolver]\ntype = pcg\ntol_rel = 1e-8\nmax_iter = 1000\nprecon = mgv\n"
                              "[mgv]\ntype = mg\ncycle = v\nhierarchy = h1\nmemorytype = hip\ndatatype = double\nformat = sell\n"
                              "[h1]\ntype = hierarchy\nsmoother = rich\ncoarse = coarse\n"
                              "[rich]\ntype = richardson\nmax_iter = 4\nmin_iter = 4\nprecon = jac  # damped Jacobi\n"
                              "[jac]\ntype = jacobi\nomega = " + std::to_string(omega) + "\n[coarse]\ntype = umfpack\n";
      std::string ini_w = ini;
      ini_w.replace(ini_w.find("cycle = v"), 9, "cycle = w");
      PropertyMap pm;
      pm.parse(ini);
      auto fs = Solver::SolverFactory::create_scalar_solver(pm, "linsolver");
      for(const HostLevel& h : host)
        fs->push_level(h.n, h.rp, h.col, h.val, h.bnd, h.nc, h.prp, h.pcol, h.pval, h.rrp, h.rcol, h.rval);
      fs->init();
      LAFEM::DenseVector<Mem::HIP, double, u32> x(h0.n, 0.0), b(rhs);
      const Solver::Status st = fs->correct(x, b);
      std::cout << "factory64 " << int(st) << " " << fs->get_num_iter() << std::endl;
      write(outdir + "/factory64.bin", x.to_host());
      // the same tree with a W-cycle (multigrid.hpp:1409-1519)
      PropertyMap pw;
      pw.parse(ini_w);
      auto fw = Solver::SolverFactory::create_scalar_solver(pw, "linsolver");
      for(const HostLevel& h : host)
        fw->push_level(h.n, h.rp, h.col, h.val, h.bnd, h.nc, h.prp, h.pcol, h.pval, h.rrp, h.rcol, h.rval);
      fw->init();
      LAFEM::DenseVector<Mem::HIP, double, u32> xw(h0.n, 0.0);
      const Solver::Status stw = fw->correct(xw, b);
      std::cout << "factoryw64 " << int(stw) << " " << fw->get_num_iter() << std::endl;
      write(outdir + "/factoryw64.bin", xw.to_host());
      bool rejected = false;
      try
      {
        PropertyMap bad;
        std::string t = ini_w;
        t.replace(t.find("cycle = w"), 9, "cycle = x");
        bad.parse(t);
        Solver::SolverFactory::create_scalar_solver(bad, "linsolver");
      }
      catch(const InternalError&)
      {
        rejected = true;
      }
      std::cout << "factory_rejects_unknown_cycle " << int(rejected) << std::endl;
    }
    {
      // the same solver over the banded level format (SparseMatrixBanded layout)
      Solver::FusedMGPCG fused(Solver::FusedMGPCG::fp64, Solver::FusedMGPCG::banded);
      for(const HostLevel& h : host)
        fused.push_level(h.n, h.rp, h.col, h.val, h.bnd, h.nc, h.prp, h.pcol, h.pval, h.rrp, h.rcol, h.rval);
      fused.set_smoother(omega, 1.0, 4, 4);
      fused.set_tol_rel(1e-8);
      fused.set_max_iter(1000);
      fused.init();
      LAFEM::DenseVector<Mem::HIP, double, u32> x(h0.n, 0.0), b(rhs);
      const Solver::Status st = fused.correct(x, b);
      std::cout << "fusedbanded64 " << int(st) << " " << fused.get_num_iter() << std::endl;
      write(outdir + "/fusedbanded64.bin", x.to_host());
    }
    {
      // SparseMatrixBanded<Mem::HIP>::apply against SparseMatrixCSR<Mem::HIP>::apply (the CSR
      // kernel reduces a row over several lanes, so the two differ by rounding only)
      LAFEM::SparseMatrixBanded<Mem::HIP, double, u32> B(h0.n, h0.n, h0.rp, h0.col, h0.val);
      std::vector<double> xs(h0.n);
      for(Index i = 0; i < h0.n; ++i) xs[i] = std::sin(0.37 * double(i)) + 0.25;
      LAFEM::DenseVector<Mem::HIP, double, u32> xv(xs), rb(h0.n, 0.0), rc(h0.n, 0.0);
      const DeviceLevel<double>::Matrix A(h0.n, h0.n, h0.rp, h0.col, h0.val);
      B.apply(rb, xv);
      A.apply(rc, xv);
      const std::vector<double> hb = rb.to_host(), hc = rc.to_host();
      double dmax = 0.0, cmax = 0.0;
      for(Index i = 0; i < h0.n; ++i)
      {
        dmax = std::max(dmax, std::abs(hb[i] - hc[i]));
        cmax = std::max(cmax, std::abs(hc[i]));
      }
      // reported in units of 1e-16 of the largest entry, as an integer
      std::cout << "banded_vs_csr " << (long long)(std::ceil(dmax / cmax * 1e16)) << " " << B.num_of_offsets() << std::endl;
    }
    {
      // the reference's file formats: write the finest matrix as fm_csr and fm_mtx (symmetric),
      // read both back into device matrices and compare their products with the original's
      const DeviceLevel<double>::Matrix A(h0.n, h0.n, h0.rp, h0.col, h0.val);
      A.write_out(LAFEM::FileMode::fm_csr, outdir + "/A.csr");
      A.write_out(LAFEM::FileMode::fm_mtx, outdir + "/A.mtx", true);
      const DeviceLevel<double>::Matrix Ab(LAFEM::FileMode::fm_csr, outdir + "/A.csr");
      const DeviceLevel<double>::Matrix Am(LAFEM::FileMode::fm_mtx, outdir + "/A.mtx");
      std::vector<double> xs(h0.n);
      for(Index i = 0; i < h0.n; ++i) xs[i] = std::cos(0.11 * double(i));
      LAFEM::DenseVector<Mem::HIP, double, u32> xv(xs), r0(h0.n, 0.0), rb(h0.n, 0.0), rm(h0.n, 0.0);
      A.apply(r0, xv);
      Ab.apply(rb, xv);
      Am.apply(rm, xv);
      const std::vector<double> h0v = r0.to_host(), hbv = rb.to_host(), hmv = rm.to_host();
      const bool same = Ab.used_elements() == A.used_elements() && Am.used_elements() == A.used_elements() &&
                        h0v == hbv && h0v == hmv;
      std::cout << "file_roundtrip " << int(same) << " " << Ab.used_elements() << std::endl;
      // vectors: fm_dv bitwise, fm_exp / fm_mtx to the 6 digits the reference writes
      r0.write_out(LAFEM::FileMode::fm_dv, outdir + "/r.dv");
      r0.write_out(LAFEM::FileMode::fm_exp, outdir + "/r.exp");
      r0.write_out(LAFEM::FileMode::fm_mtx, outdir + "/r.mtx");
      const LAFEM::DenseVector<Mem::HIP, double, u32> vd(LAFEM::FileMode::fm_dv, outdir + "/r.dv");
      const LAFEM::DenseVector<Mem::HIP, double, u32> ve(LAFEM::FileMode::fm_exp, outdir + "/r.exp");
      LAFEM::DenseVector<Mem::HIP, double, u32> vm;
      vm.read_from(LAFEM::FileMode::fm_mtx, outdir + "/r.mtx");
      const std::vector<double> hd = vd.to_host(), he = ve.to_host(), hm = vm.to_host();
      bool vsame = hd == h0v && he.size() == h0v.size() && hm.size() == h0v.size();
      for(std::size_t i = 0; vsame && i < h0v.size(); ++i)
        vsame = std::abs(he[i] - h0v[i]) <= 5e-7 * std::abs(h0v[i]) && he[i] == hm[i];
      std::cout << "vector_files " << int(vsame) << " " << vd.size() << std::endl;
    }
    std::cout << "leaked " << MemoryPool<Mem::HIP>::finalise() << std::endl;
  }
  catch(const std::exception& e)
  {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
