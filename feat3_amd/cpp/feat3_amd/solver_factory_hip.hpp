// solver_factory_hip.hpp -- SolverFactory for the device PCG-MG path (the reference's
// kernel/solver/solver_factory.hpp:376-640 with its PropertyMap INI configuration,
// data/ini/solver_example.ini).
//
// PropertyMap holds [section] blocks of "key = value" lines ('#' starts a comment).
// SolverFactory::create_scalar_solver(base, "linsolver") follows the tree from the named root
// and configures the fused solver for the shape it implements:
//
//   root:      type = pcg                      tol_rel, tol_abs, tol_abs_low, max_iter, min_iter
//   precon:    type = mg, cycle = v | f | w    hierarchy, lvl_min = -1, lvl_max = 0,
//                                              memorytype = hip, datatype = float | double
//                                              (+ valuetype = half: binary16 matrix values)
//   hierarchy: type = hierarchy                smoother, coarse
//   smoother:  type = richardson               min_iter = max_iter = steps, omega,
//                                              precon = <type = jacobi, omega>
//              or type = jacobi                (one damped Jacobi step)
//   coarse:    type = umfpack | cusolver | direct   (exact FP64 solve)
//
// "precon = none" gives plain CG.  Anything else throws InternalError naming the section; the
// extension keys "format = csr | sell | banded" and "graph = 0 | 1" select the level format and
// graph replay.  The caller pushes the levels and calls init() as for a directly built solver.
#pragma once

#include "solver_hip.hpp"

#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>

namespace feat3_amd
{
  class PropertyMap
  {
  public:
    typedef std::map<std::string, std::string> Section;

    static std::string trim(const std::string& s)
    {
      const std::size_t b = s.find_first_not_of(" \t\r");
      if(b == std::string::npos) return std::string();
      const std::size_t e = s.find_last_not_of(" \t\r");
      return s.substr(b, e - b + 1);
    }

    // parses INI text (lines before the first section go to section "")
    void parse(const std::string& text)
    {
      std::istringstream is(text);
      std::string raw, cur;
      while(std::getline(is, raw))
      {
        const std::string line = trim(raw.substr(0, raw.find('#')));
        if(line.empty()) continue;
        if(line.front() == '[' && line.back() == ']')
        {
          cur = trim(line.substr(1, line.size() - 2));
          _s[cur];
          continue;
        }
        const std::size_t eq = line.find('=');
        if(eq == std::string::npos) throw InternalError("PropertyMap: expected 'key = value' or '[section]': " + raw);
        _s[cur][trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
      }
    }

    void read(const std::string& filename)
    {
      std::ifstream f(filename);
      if(!f.is_open()) throw InternalError("PropertyMap: cannot open " + filename);
      std::stringstream ss;
      ss << f.rdbuf();
      parse(ss.str());
    }

    const Section* query_section(const std::string& name) const
    {
      auto it = _s.find(name);
      return it == _s.end() ? nullptr : &it->second;
    }

  private:
    std::map<std::string, Section> _s;
  };

  namespace Solver
  {
    class SolverFactory
    {
      typedef PropertyMap::Section Section;

      static const Section& section(const PropertyMap& base, const std::string& name, const std::string& parent)
      {
        const Section* s = base.query_section(name);
        if(!s) throw InternalError("section not found in property map: " + name + " (referenced by " + parent + ")");
        return *s;
      }
      static bool has(const Section& s, const char* key) { return s.find(key) != s.end(); }
      static std::string get(const Section& s, const char* key, const std::string& def)
      {
        auto it = s.find(key);
        return it == s.end() ? def : it->second;
      }
      static double num(const Section& s, const std::string& where, const char* key, double def)
      {
        auto it = s.find(key);
        if(it == s.end()) return def;
        char* end = nullptr;
        const double v = std::strtod(it->second.c_str(), &end);
        if(end == it->second.c_str() || *end != '\0') throw InternalError(where + "." + key + ": cannot parse " + it->second);
        return v;
      }
      static Index count(const Section& s, const std::string& where, const char* key, Index def)
      {
        const double v = num(s, where, key, double(def));
        if(v < 0.0 || v != double(Index(v))) throw InternalError(where + "." + key + ": an integer >= 0 expected");
        return Index(v);
      }

    public:
      static std::shared_ptr<FusedMGPCG> create_scalar_solver(const PropertyMap& base, const std::string& name)
      {
        const Section& root = section(base, name, "the caller");
        if(get(root, "type", "") != "pcg")
          throw InternalError(name + ".type = " + get(root, "type", "") + ": this path is PCG preconditioned by multigrid");
        if(get(root, "datatype", "double") != "double") throw InternalError(name + ".datatype: the outer PCG runs in double");
        for(const char* k : {"div_rel", "div_abs", "stag_rate", "min_stag_iter"})
          if(has(root, k)) throw InternalError(name + "." + k + ": not configurable on the device PCG");
        std::string fmt = get(root, "format", "sell");
        FusedMGPCG::Inner inner = FusedMGPCG::fp64;
        const std::string pre = get(root, "precon", "none");
        double oj = 0.5, orr = 1.0;
        int steps = 4;
        MultiGridCycle cycle = MultiGridCycle::V;
        if(pre != "none")
        {
          const Section& mg = section(base, pre, name + ".precon");
          if(get(mg, "type", "") != "mg") throw InternalError(pre + ".type: the preconditioner must be a multigrid (type = mg)");
          if(!has(mg, "cycle")) throw InternalError(pre + ": mg section without cycle key is not allowed");
          const std::string cyc = get(mg, "cycle", "");
          if(cyc == "v" || cyc == "V") cycle = MultiGridCycle::V;
          else if(cyc == "f" || cyc == "F") cycle = MultiGridCycle::F;
          else if(cyc == "w" || cyc == "W") cycle = MultiGridCycle::W;
          else throw InternalError(pre + ".cycle = " + cyc + ": mg cycle unknown (v, f or w)");
          if(num(mg, pre, "lvl_max", 0.0) != 0.0 || num(mg, pre, "lvl_min", -1.0) != -1.0)
            throw InternalError(pre + ": the cycle spans the whole hierarchy (lvl_max = 0, lvl_min = -1)");
          if(get(mg, "adapt_cgc", "fixed") != "fixed") throw InternalError(pre + ".adapt_cgc: the correction is fixed");
          const std::string mem = get(mg, "memorytype", "hip");
          if(mem != "hip" && mem != "main") throw InternalError(pre + ".memorytype = " + mem + ": the cycle runs on the device");
          const std::string dt = get(mg, "datatype", "double"), vt = get(mg, "valuetype", "");
          if(vt == "half") inner = FusedMGPCG::fp16;
          else if(!vt.empty()) throw InternalError(pre + ".valuetype = " + vt + ": half or absent");
          else if(dt == "float") inner = FusedMGPCG::fp32;
          else if(dt != "double") throw InternalError(pre + ".datatype = " + dt + ": float or double");
          fmt = get(mg, "format", fmt);
          if(!has(mg, "hierarchy")) throw InternalError(pre + ": no hierarchy key found in mg section");
          const std::string hn = get(mg, "hierarchy", "");
          const Section& h = section(base, hn, pre + ".hierarchy");
          if(get(h, "type", "") != "hierarchy") throw InternalError(hn + ".type must be 'hierarchy'");
          if(!has(h, "smoother") || !has(h, "coarse")) throw InternalError(hn + ": a hierarchy needs smoother and coarse keys");
          const std::string sn = get(h, "smoother", "");
          const Section& sm = section(base, sn, hn + ".smoother");
          const std::string st = get(sm, "type", "");
          if(st == "jacobi")
          {
            oj = num(sm, sn, "omega", 1.0);
            orr = 1.0;
            steps = 1;
          }
          else if(st == "richardson")
          {
            const Index mx = count(sm, sn, "max_iter", 100), mn = count(sm, sn, "min_iter", 0);
            if(mx != mn) throw InternalError(sn + ": a smoother runs a fixed number of steps here (min_iter = max_iter)");
            if(mx < 1) throw InternalError(sn + ".max_iter: at least one smoothing step");
            if(count(sm, sn, "min_stag_iter", 0) != 0) throw InternalError(sn + ".min_stag_iter: no stagnation test");
            orr = num(sm, sn, "omega", 1.0);
            const std::string pj = get(sm, "precon", "none");
            if(pj == "none") throw InternalError(sn + ".precon: the device smoother is richardson with a jacobi preconditioner");
            const Section& js = section(base, pj, sn + ".precon");
            if(get(js, "type", "") != "jacobi") throw InternalError(pj + ".type: the smoother's preconditioner must be jacobi");
            oj = num(js, pj, "omega", 1.0);
            steps = int(mx);
          }
          else
            throw InternalError(sn + ".type = " + st + ": the device multigrid smooths with richardson(jacobi) or jacobi");
          const std::string cn = get(h, "coarse", "");
          const std::string ct = get(section(base, cn, hn + ".coarse"), "type", "");
          if(ct != "umfpack" && ct != "cusolver" && ct != "direct")
            throw InternalError(cn + ".type = " + ct + ": the coarse level is solved exactly (umfpack | cusolver | direct)");
        }
        FusedMGPCG::Format f = FusedMGPCG::sell;
        if(fmt == "csr") f = FusedMGPCG::csr;
        else if(fmt == "banded") f = FusedMGPCG::banded;
        else if(fmt != "sell") throw InternalError("format = " + fmt + ": one of csr, sell, banded");
        auto solver = std::make_shared<FusedMGPCG>(inner, f);
        solver->set_smoother(oj, orr, steps, steps);
        if(pre == "none") solver->set_precond_none();
        else if(cycle != MultiGridCycle::V) solver->set_cycle(cycle);
        if(has(root, "tol_rel")) solver->set_tol_rel(num(root, name, "tol_rel", 0.0));
        if(has(root, "tol_abs")) solver->set_tol_abs(num(root, name, "tol_abs", 0.0));
        if(has(root, "tol_abs_low")) solver->set_tol_abs_low(num(root, name, "tol_abs_low", 0.0));
        if(has(root, "max_iter")) solver->set_max_iter(count(root, name, "max_iter", 0));
        if(has(root, "min_iter")) solver->set_min_iter(count(root, name, "min_iter", 0));
        if(has(root, "graph")) solver->set_graph(num(root, name, "graph", 1.0) != 0.0);
        return solver;
      }
    };
  } // namespace Solver
} // namespace feat3_amd
