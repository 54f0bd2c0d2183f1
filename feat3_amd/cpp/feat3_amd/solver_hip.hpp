// solver_hip.hpp -- the Solver side of the reference interface for Mem::HIP containers.
//
// Two ways to run PCG( MultiGrid-V( Richardson(Jacobi) ) ) on the device, both with the
// reference's SolverBase lifecycle (init_symbolic / init_numeric / apply / correct / done,
// kernel/solver/base.hpp:184-316):
//
//  * the generic, op-by-op classes PCG, Richardson, JacobiPrecond, MultiGrid(+Hierarchy) and
//    ConvertPrecond below.  They restate the reference algorithms on top of the container
//    methods (each reduction returns to the host, as the reference's CUDA backend does), so
//    any LAFEM-style matrix/filter/vector triple runs unchanged on the HIP tag:
//      PCG          kernel/solver/pcg.hpp:140-288
//      criteria     kernel/solver/iterative.hpp:213-237,476-505,744-879,1063-1076
//      Richardson   kernel/solver/richardson.hpp:160-228
//      Jacobi       kernel/solver/jacobi_precond.hpp:137-165
//      MultiGrid    kernel/solver/multigrid.hpp:419-957,1268-1351,1524-1587,1711-2000
//      Convert      kernel/solver/convert_precond.hpp:100-119
//  * FusedMGPCG: the same solver as one device-resident object (fused smoothing kernels,
//    device scalars, hipGraph replay) behind the feat3_mgpcg_* C ABI -- the fast path.
#pragma once

#include "lafem_hip.hpp"

#include <memory>
#include <vector>

namespace feat3_amd
{
  namespace Solver
  {
    // base.hpp:48-64
    enum class Status { undefined = 0, progress, success, aborted, diverged, max_iter, stagnated };

    // base.hpp:110-123
    inline bool status_success(Status s)
    {
      return s == Status::success || s == Status::max_iter || s == Status::stagnated;
    }

    template<typename Vector_>
    class SolverBase
    {
    public:
      typedef Vector_ VectorType;
      virtual ~SolverBase() {}
      virtual void init_symbolic() {}
      virtual void init_numeric() {}
      virtual void init() { init_symbolic(); init_numeric(); }
      virtual void done_numeric() {}
      virtual void done_symbolic() {}
      virtual void done() { done_numeric(); done_symbolic(); }
      virtual Status apply(Vector_& vec_cor, const Vector_& vec_def) = 0;
    };

    // iterative.hpp:55-1077
    template<typename Vector_>
    class IterativeSolver : public SolverBase<Vector_>
    {
    public:
      typedef typename Vector_::DataType DataType;
    protected:
      DataType _tol_rel = std::sqrt(Math::eps<DataType>());
      DataType _tol_abs = DataType(1) / (Math::eps<DataType>() * Math::eps<DataType>());
      DataType _tol_abs_low = DataType(0);
      DataType _div_rel = DataType(1) / Math::eps<DataType>();
      DataType _div_abs = DataType(1) / (Math::eps<DataType>() * Math::eps<DataType>());
      DataType _stag_rate = DataType(0.95);
      Index _min_iter = 0, _max_iter = 100, _num_iter = 0, _min_stag_iter = 0, _num_stag_iter = 0;
      DataType _def_init = 0, _def_cur = 0, _def_prev = 0;
      bool _skip_def_calc = true;
      Status _status = Status::undefined;
      std::shared_ptr<SolverBase<Vector_>> _precond;
    public:
      void set_tol_rel(DataType t) { _tol_rel = t; }
      void set_tol_abs(DataType t) { _tol_abs = t; }
      void set_tol_abs_low(DataType t) { _tol_abs_low = t; }
      void set_min_iter(Index i) { _min_iter = i; }
      void set_max_iter(Index i) { _max_iter = i; }
      Index get_num_iter() const { return _num_iter; }
      DataType get_def_initial() const { return _def_init; }
      DataType get_def_final() const { return _def_cur; }
      Status get_status() const { return _status; }
      bool is_converged(DataType d) const { return (d <= _tol_abs) && ((d <= _tol_rel * _def_init) || (d <= _tol_abs_low)); }
      bool is_diverged(DataType d) const { return (d > _div_abs) || (d > _div_rel * _def_init); }
      void init_symbolic() override { if(_precond) _precond->init_symbolic(); }
      void init_numeric() override { if(_precond) _precond->init_numeric(); }
      virtual Status correct(Vector_& vec_sol, const Vector_& vec_rhs) = 0;
    protected:
      Status _set_initial_defect(const Vector_& vec_def)
      {
        _def_init = _def_cur = _def_prev = vec_def.norm2();
        _num_iter = 0;
        _num_stag_iter = 0;
        if(!std::isfinite(double(_def_init))) return Status::aborted;
        if(_def_init <= Math::eps<DataType>() * Math::eps<DataType>()) return Status::success;
        return Status::progress;
      }
      Status _analyse_defect(Index it, DataType dc, DataType dp)
      {
        if(!std::isfinite(double(dc))) return Status::aborted;
        if(is_diverged(dc)) return Status::diverged;
        if(it < _min_iter) return Status::progress;
        if(is_converged(dc)) return Status::success;
        if(it >= _max_iter) return Status::max_iter;
        if(_min_stag_iter > 0)
        {
          if(dc >= _stag_rate * dp) { if(++_num_stag_iter >= _min_stag_iter) return Status::stagnated; }
          else _num_stag_iter = 0;
        }
        return Status::progress;
      }
      Status _set_new_defect(const Vector_& vec_def)
      {
        ++_num_iter;
        _def_prev = _def_cur;
        if(!_skip_def_calc || (_min_iter < _max_iter) || (_min_stag_iter > 0))
          _def_cur = vec_def.norm2();
        return _analyse_defect(_num_iter, _def_cur, _def_prev);
      }
      template<typename Filter_>
      bool _apply_precond(Vector_& cor, const Vector_& def, const Filter_& filter)
      {
        if(_precond) return status_success(_precond->apply(cor, def));
        cor.copy(def);
        filter.filter_cor(cor);
        return true;
      }
    };

    // jacobi_precond.hpp:36-213
    template<typename Matrix_, typename Filter_>
    class JacobiPrecond : public SolverBase<typename Matrix_::VectorTypeL>
    {
      typedef typename Matrix_::VectorTypeL VectorType;
      typedef typename VectorType::DataType DataType;
      const Matrix_& _matrix;
      const Filter_& _filter;
      DataType _omega;
      VectorType _inv_diag;
    public:
      JacobiPrecond(const Matrix_& m, const Filter_& f, DataType omega = DataType(1)) : _matrix(m), _filter(f), _omega(omega) {}
      void init_symbolic() override { _inv_diag = _matrix.create_vector_r(); }
      void init_numeric() override
      {
        _matrix.extract_diag(_inv_diag);
        _inv_diag.component_invert(_inv_diag, _omega);
      }
      Status apply(VectorType& cor, const VectorType& def) override
      {
        cor.component_product(_inv_diag, def);
        _filter.filter_cor(cor);
        return Status::success;
      }
    };

    // richardson.hpp:32-230
    template<typename Matrix_, typename Filter_>
    class Richardson : public IterativeSolver<typename Matrix_::VectorTypeL>
    {
      typedef typename Matrix_::VectorTypeL VectorType;
      typedef typename VectorType::DataType DataType;
      const Matrix_& _matrix;
      const Filter_& _filter;
      DataType _omega;
      VectorType _vec_def, _vec_cor;
    public:
      Richardson(const Matrix_& m, const Filter_& f, DataType omega, std::shared_ptr<SolverBase<VectorType>> precond)
        : _matrix(m), _filter(f), _omega(omega) { this->_precond = precond; }
      void init_symbolic() override
      {
        IterativeSolver<VectorType>::init_symbolic();
        _vec_def = _matrix.create_vector_r();
        _vec_cor = _matrix.create_vector_r();
      }
      Status apply(VectorType& cor, const VectorType& def) override
      {
        _vec_def.copy(def);
        cor.format();
        return this->_status = _apply_intern(cor, def);
      }
      Status correct(VectorType& sol, const VectorType& rhs) override
      {
        _matrix.apply(_vec_def, sol, rhs, -DataType(1));
        _filter.filter_def(_vec_def);
        return this->_status = _apply_intern(sol, rhs);
      }
    private:
      Status _apply_intern(VectorType& sol, const VectorType& rhs)
      {
        Status st = this->_set_initial_defect(_vec_def);
        while(st == Status::progress)
        {
          if(!this->_apply_precond(_vec_cor, _vec_def, _filter)) return Status::aborted;
          sol.axpy(_vec_cor, sol, _omega);
          _matrix.apply(_vec_def, sol, rhs, -DataType(1));
          _filter.filter_def(_vec_def);
          st = this->_set_new_defect(_vec_def);
        }
        return st;
      }
    };

    // pcg.hpp:33-328
    template<typename Matrix_, typename Filter_>
    class PCG : public IterativeSolver<typename Matrix_::VectorTypeL>
    {
      typedef typename Matrix_::VectorTypeL VectorType;
      typedef typename VectorType::DataType DataType;
      const Matrix_& _matrix;
      const Filter_& _filter;
      VectorType _vec_r, _vec_p, _vec_t;
    public:
      PCG(const Matrix_& m, const Filter_& f, std::shared_ptr<SolverBase<VectorType>> precond = nullptr)
        : _matrix(m), _filter(f) { this->_precond = precond; }
      void init_symbolic() override
      {
        IterativeSolver<VectorType>::init_symbolic();
        _vec_r = _matrix.create_vector_r();
        _vec_p = _matrix.create_vector_r();
        _vec_t = _matrix.create_vector_r();
      }
      Status apply(VectorType& cor, const VectorType& def) override
      {
        _vec_r.copy(def);
        cor.format();
        return this->_status = _apply_intern(cor);
      }
      Status correct(VectorType& sol, const VectorType& rhs) override
      {
        _matrix.apply(_vec_r, sol, rhs, -DataType(1));
        _filter.filter_def(_vec_r);
        return this->_status = _apply_intern(sol);
      }
    private:
      Status _apply_intern(VectorType& sol)
      {
        VectorType& q = _vec_t;
        VectorType& z = _vec_t;
        Status st = this->_set_initial_defect(_vec_r);
        if(st != Status::progress) return st;
        if(!this->_apply_precond(_vec_p, _vec_r, _filter)) return Status::aborted;
        DataType gamma = _vec_r.dot(_vec_p);
        while(st == Status::progress)
        {
          _matrix.apply(q, _vec_p);
          _filter.filter_def(q);
          const DataType alpha = gamma / q.dot(_vec_p);
          sol.axpy(_vec_p, sol, alpha);
          _vec_r.axpy(q, _vec_r, -alpha);
          st = this->_set_new_defect(_vec_r);
          if(st != Status::progress) return st;
          if(!this->_apply_precond(z, _vec_r, _filter)) return Status::aborted;
          const DataType gamma2 = gamma;
          gamma = _vec_r.dot(z);
          _vec_p.axpy(_vec_p, z, gamma / gamma2);
        }
        return Status::undefined;
      }
    };

    // multigrid.hpp:419-957 (hierarchy) and 1268-2000 (V-cycle with fixed coarse-grid correction)
    template<typename Matrix_, typename Filter_, typename Transfer_>
    class MultiGridHierarchy
    {
    public:
      typedef typename Matrix_::VectorTypeL VectorType;
      struct Level
      {
        const Matrix_* matrix;
        const Filter_* filter;
        const Transfer_* transfer;   // to the next coarser level (nullptr on the coarsest)
        std::shared_ptr<SolverBase<VectorType>> smoother_pre, smoother_post, coarse_solver;
        VectorType rhs, sol, def, cor, tmp;
      };
      std::vector<Level> levels;   // finest first (push_level order of the reference)

      void push_level(const Matrix_& m, const Filter_& f, const Transfer_& t, std::shared_ptr<SolverBase<VectorType>> pre,
                      std::shared_ptr<SolverBase<VectorType>> post, std::shared_ptr<SolverBase<VectorType>> coarse = nullptr)
      {
        levels.push_back(Level{&m, &f, &t, pre, post, coarse, {}, {}, {}, {}, {}});
      }
      void push_level(const Matrix_& m, const Filter_& f, std::shared_ptr<SolverBase<VectorType>> coarse)
      {
        levels.push_back(Level{&m, &f, nullptr, nullptr, nullptr, coarse, {}, {}, {}, {}, {}});
      }
      void init()
      {
        for(Level& l : levels)
        {
          l.rhs = l.matrix->create_vector_r(); l.sol = l.matrix->create_vector_r(); l.def = l.matrix->create_vector_r();
          l.cor = l.matrix->create_vector_r(); l.tmp = l.matrix->create_vector_r();
          for(auto* s : {&l.smoother_pre, &l.smoother_post, &l.coarse_solver})
            if(*s) (*s)->init();
        }
      }
    };

    template<typename Matrix_, typename Filter_, typename Transfer_>
    class MultiGrid : public SolverBase<typename Matrix_::VectorTypeL>
    {
      typedef typename Matrix_::VectorTypeL VectorType;
      typedef typename VectorType::DataType DataType;
      std::shared_ptr<MultiGridHierarchy<Matrix_, Filter_, Transfer_>> _h;
    public:
      explicit MultiGrid(std::shared_ptr<MultiGridHierarchy<Matrix_, Filter_, Transfer_>> h) : _h(h) {}
      Status apply(VectorType& vec_cor, const VectorType& vec_def) override
      {
        auto& lv = _h->levels;
        lv.front().rhs.copy(vec_def);
        // _apply_rest (multigrid.hpp:1711-1832)
        for(std::size_t i = 0; i + 1 < lv.size(); ++i)
        {
          auto& f = lv[i];
          if(f.smoother_pre)
          {
            f.smoother_pre->apply(f.sol, f.rhs);
            f.matrix->apply(f.def, f.sol, f.rhs, -DataType(1));
          }
          else
          {
            f.sol.format();
            f.def.copy(f.rhs);
          }
          f.filter->filter_def(f.def);
          f.transfer->rest(f.def, lv[i + 1].rhs);
          lv[i + 1].filter->filter_def(lv[i + 1].rhs);
        }
        // _apply_coarse (multigrid.hpp:1524-1587)
        auto& c = lv.back();
        if(c.coarse_solver)
        {
          if(!status_success(c.coarse_solver->apply(c.sol, c.rhs))) return Status::aborted;
        }
        else
        {
          c.sol.copy(c.rhs);
          c.filter->filter_cor(c.sol);
        }
        // _apply_prol (multigrid.hpp:1843-2000)
        for(std::size_t ii = lv.size() - 1; ii > 0; --ii)
        {
          auto& f = lv[ii - 1];
          f.transfer->prol(f.cor, lv[ii].sol);
          f.filter->filter_cor(f.cor);
          f.sol.axpy(f.cor, f.sol, DataType(1));
          if(f.smoother_post)
          {
            f.matrix->apply(f.def, f.sol, f.rhs, -DataType(1));
            f.filter->filter_def(f.def);
            f.smoother_post->apply(f.cor, f.def);
            f.sol.axpy(f.cor, f.sol, DataType(1));
          }
        }
        vec_cor.copy(lv.front().sol);
        return Status::success;
      }
    };

    // convert_precond.hpp:41-237 (outer vector type <- inner solver in another precision)
    template<typename VectorOuter_, typename VectorInner_>
    class ConvertPrecond : public SolverBase<VectorOuter_>
    {
      std::shared_ptr<SolverBase<VectorInner_>> _inner;
      VectorInner_ _def, _cor;
    public:
      explicit ConvertPrecond(std::shared_ptr<SolverBase<VectorInner_>> inner) : _inner(inner) {}
      void init_symbolic() override { _inner->init_symbolic(); }
      void init_numeric() override { _inner->init_numeric(); }
      Status apply(VectorOuter_& cor, const VectorOuter_& def) override
      {
        _def.convert(def);
        if(_cor.size() != _def.size()) _cor = _def.clone(LAFEM::CloneMode::Layout);
        const Status st = _inner->apply(_cor, _def);
        if(!status_success(st)) return st;
        cor.convert(_cor);
        return Status::success;
      }
    };

    // Solver::solve for iterative solvers (iterative.hpp:926-935) and for the rest (base.hpp:343-375)
    template<typename Vector_, typename Matrix_, typename Filter_>
    Status solve(IterativeSolver<Vector_>& solver, Vector_& sol, const Vector_& rhs, const Matrix_&, const Filter_&)
    {
      return solver.correct(sol, rhs);
    }

    // ---- the fused device-resident solver ---------------------------------------------------
    //
    // PCG( MultiGrid-V( Richardson(omega_R, Jacobi(omega_J)) x pre/post ), coarse = exact FP64 ),
    // FP64 outer iteration, multigrid cycle (V, F or W) in FP64 / FP32 / FP16-stored values.  The level data is
    // given as host CSR arrays (finest first), exactly what the hierarchy above holds.
    // multigrid cycle kinds (multigrid.hpp:26-34)
    enum class MultiGridCycle { V, F, W };

    class FusedMGPCG : public IterativeSolver<LAFEM::DenseVector<Mem::HIP, double, std::uint32_t>>
    {
      feat3_mgpcg* _h = nullptr;
    public:
      typedef LAFEM::DenseVector<Mem::HIP, double, std::uint32_t> VectorType;
      enum Inner { fp64 = 0, fp32 = 1, fp16 = 2 };
      enum Format { csr = 0, sell = 1, banded = 2 };

      FusedMGPCG(Inner inner = fp64, Format fmt = sell) { hip_check(feat3_mgpcg_create(&_h, int(inner), int(fmt)), "FusedMGPCG"); }
      ~FusedMGPCG() override { if(_h) feat3_mgpcg_destroy(_h); }
      FusedMGPCG(const FusedMGPCG&) = delete;
      FusedMGPCG& operator=(const FusedMGPCG&) = delete;

      void push_level(Index n, const std::vector<std::uint32_t>& rp, const std::vector<std::uint32_t>& col,
                      const std::vector<double>& val, const std::vector<std::uint32_t>& dirichlet,
                      Index nc = 0, const std::vector<std::uint32_t>& prp = {}, const std::vector<std::uint32_t>& pcol = {},
                      const std::vector<double>& pval = {}, const std::vector<std::uint32_t>& rrp = {},
                      const std::vector<std::uint32_t>& rcol = {}, const std::vector<double>& rval = {})
      {
        hip_check(feat3_mgpcg_push_level(_h, n, rp.data(), col.data(), val.data(), dirichlet.size(), dirichlet.data(), nc,
                                         nc ? prp.data() : nullptr, nc ? pcol.data() : nullptr, nc ? pval.data() : nullptr,
                                         nc ? rrp.data() : nullptr, nc ? rcol.data() : nullptr, nc ? rval.data() : nullptr),
                  "FusedMGPCG::push_level");
      }
      void set_smoother(double omega_jacobi, double omega_richardson, int pre, int post)
      {
        hip_check(feat3_mgpcg_set_smoother(_h, omega_jacobi, omega_richardson, pre, post), "FusedMGPCG::set_smoother");
      }
      // plain CG (no preconditioner) instead of the multigrid cycle
      void set_precond_none() { hip_check(feat3_mgpcg_set_precond(_h, 1), "FusedMGPCG::set_precond_none"); }
      // the multigrid cycle (default V); takes effect at the next init
      void set_cycle(MultiGridCycle c) { set_option("cycle", c == MultiGridCycle::V ? 0.0 : c == MultiGridCycle::F ? 1.0 : 2.0); }
      // graph replay of the iteration (default on)
      void set_graph(bool on) { hip_check(feat3_mgpcg_set_launch(_h, on ? 1 : 0, 0, 64, 1), "FusedMGPCG::set_graph"); }
      // a named launch/format option of feat3_mgpcg_set_option
      void set_option(const std::string& name, double value)
      {
        hip_check(feat3_mgpcg_set_option(_h, name.c_str(), value), "FusedMGPCG::set_option");
      }
      void init_numeric() override
      {
        hip_check(feat3_mgpcg_set_tolerances(_h, _tol_rel, _tol_abs, _tol_abs_low, _min_iter, _max_iter), "FusedMGPCG::init");
        hip_check(feat3_mgpcg_init(_h), "FusedMGPCG::init");
      }
      Status apply(VectorType& cor, const VectorType& def) override
      {
        int st = 0;
        std::uint64_t it = 0;
        hip_check(feat3_mgpcg_apply(_h, nullptr, cor.elements(), def.elements(), &st, &it), "FusedMGPCG::apply");
        _num_iter = it;
        return _status = Status(st);
      }
      Status correct(VectorType& sol, const VectorType& rhs) override
      {
        int st = 0;
        std::uint64_t it = 0;
        hip_check(feat3_mgpcg_correct(_h, nullptr, sol.elements(), rhs.elements(), &st, &it), "FusedMGPCG::correct");
        _num_iter = it;
        return _status = Status(st);
      }
    };
  } // namespace Solver
} // namespace feat3_amd
