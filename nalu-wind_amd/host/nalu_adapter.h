// nalu_adapter.h -- C++ host classes that mirror the reference's
// solver-algorithm / linear-system interface for the edge path and forward
// to the C ABI of libnaluwind_amd (include/nalu_amd.h).  They show how the
// reference's call sites map onto the library:
//
//   reference (Exawind/nalu-wind)                      here
//   SolverAlgorithm::execute / initialize_connectivity  same names
//     (include/SolverAlgorithm.h:85-86)
//   AssembleEdgeSolverAlgorithm                         same name; one launch
//     (include/AssembleEdgeSolverAlgorithm.h:31-109)    per execute()
//   ScalarEdgeSolverAlg / MomentumEdgeSolverAlg /       same names
//   ContinuityEdgeSolverAlg, MomentumEdgePecletAlg
//     (src/edge_kernels/*.C)
//   FixPressureAtNodeAlgorithm                          same name
//     (src/FixPressureAtNodeAlgorithm.C:57-124)
//   MdotEdgeAlg, NodalGradEdgeAlg, GeometryInteriorAlg  same names
//     (src/ngp_algorithms/{MdotEdgeAlg,NodalGradEdgeAlg,GeometryInteriorAlg}.C)
//   LinearSystem / TpetraLinearSystem                   LinearSystem /
//     (include/LinearSystem.h:44-256)                   GpuLinearSystem
//   CoeffApplier (include/LinearSystem.h:44-74)         CoeffApplier
//   EquationSystem::assemble_and_solve                  EquationSystem::
//     (src/EquationSystem.C:301-350)                    assemble_and_solve
//   Realm::create_mesh / populate_mesh / create_edges   Realm::create_mesh
//     (src/Realm.C:2048-2094, 450-466), the mesh file   (Exodus II, netCDF
//     of a boundary condition's target_name             classic), bc_nodes
//   Realm::create_output_mesh / output              Realm::write_output
//     (src/Realm.C:2100-2215)
//
// Errors from the library become std::runtime_error, as the reference's
// host code throws; nothing is thrown across the C ABI itself.
#pragma once

#include <chrono>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "nalu_amd.h"
#include "nalu_amd_mesh.h"

namespace nalu_amd {

void check(int rc, const char* what);

// Scheme options the edge algorithms read (SolutionOptions defaults,
// src/SolutionOptions.C:32-52), per degree-of-freedom name.
struct SolutionOptions {
  double hybridDefault = 0.0, alphaDefault = 0.0, alphaUpwDefault = 1.0,
         upwDefault = 1.0, relaxDefault = 1.0;
  std::string tanhFormDefault = "classic";
  double tanhTransDefault = 2.0, tanhWidthDefault = 4.0;
  bool nocDefault = true;
  double includeDivU = 0.0;
  bool mdotInterpRhoUTogether = true;
  bool solveIncompressibleContinuity = false;
  // use_balanced_buoyancy_force_ and the gravity vector
  // (include/SolutionOptions.h; read by ContinuityEdgeSolverAlg and
  // MdotEdgeAlg, which then also read the "buoyancy_source" and
  // "buoyancy_source_mask" node fields)
  bool use_balanced_buoyancy_force_ = false;
  double gravity[3] = {0.0, 0.0, 0.0};
  std::map<std::string, double> hybridMap, alphaMap, alphaUpwMap, upwMap,
    relaxMap;
  std::map<std::string, bool> limiterMap, nocMap;

  double get(const std::map<std::string, double>& m, const std::string& k,
             double d) const
  {
    auto it = m.find(k);
    return it == m.end() ? d : it->second;
  }
  nw_peclet peclet(const std::string& dof) const;
};

struct TimeIntegrator {
  double timeStepN = 1.0;
  double gamma1 = 1.0;
};

class LinearSystem;

// The part of Realm the edge path needs: one GPU context holding the mesh
// (nodes, edges) and its fields, the options and the time integrator.
//
// The context holds one linear system at a time.  The reference gives each
// equation system of a Realm its own linear system; here the systems of
// one Realm take turns: the one whose graph the context holds is the
// active one, zeroSystem() of another system makes that one active
// (rebuilding its graph, which also drops the previous system's matrix,
// solver workspace and preconditioner), and every other call on an
// inactive system throws.  Assembling several systems therefore costs one
// graph build per switch; a caller that keeps several systems resident
// gives each its own Realm.
class Realm {
 public:
  explicit Realm(int device = 0);
  ~Realm();
  Realm(const Realm&) = delete;
  Realm& operator=(const Realm&) = delete;

  void set_mesh(int64_t nnode, const std::vector<int32_t>& edge_nodes);
  // create_mesh + populate_mesh + create_edges for an Exodus II file
  // (netCDF classic or 64-bit-offset form): the nodes, the edges of its
  // Hex8/Tet4/Wed6/Pyr5 elements and the coordinates go to the context;
  // the elements stay with the realm for GeometryInteriorAlg(realm).
  void create_mesh(const std::string& path);
  // The nodes of the node set, or else of the side set, named `name` in
  // the mesh file (a boundary condition's target_name), sorted and unique.
  std::vector<int32_t> bc_nodes(const std::string& name) const;
  // The mesh and node fields at `time` to an Exodus II file: `fields`
  // pairs an output name with a context field; a 3-component field is
  // written as name_x, name_y, name_z.
  void write_output(const std::string& path,
                    const std::vector<std::pair<std::string, int>>& fields,
                    double time) const;
  int64_t nelem() const { return (int64_t)elemType_.size(); }
  const std::vector<int32_t>& elemTypes() const { return elemType_; }
  const std::vector<int32_t>& elemConn() const { return elemConn_; }
  void set_field(int field, const std::vector<double>& data);
  std::vector<double> get_field(int field) const;
  int64_t nnode() const { return nnode_; }
  int64_t nedge() const { return nedge_; }
  nw_ctx* ctx() const { return ctx_; }
  // the linear system whose graph the context holds (null: none)
  const LinearSystem* activeSystem() const { return active_; }
  void setActiveSystem(const LinearSystem* ls) { active_ = ls; }

  SolutionOptions solutionOptions;
  TimeIntegrator timeIntegrator;

 private:
  nw_ctx* ctx_ = nullptr;
  int64_t nnode_ = 0, nedge_ = 0;
  const LinearSystem* active_ = nullptr;
  std::string meshPath_;
  std::vector<int32_t> elemType_, elemConn_;  // [nelem], [nelem][8]
};

// LinearSystem::CoeffApplier: additive block scatter + row reset.
class CoeffApplier {
 public:
  virtual ~CoeffApplier() = default;
  virtual void operator()(unsigned numEntities, const int32_t* entities,
                          const double* rhs, const double* lhs) = 0;
  virtual void resetRows(unsigned numNodes, const int32_t* nodeList,
                         unsigned beginPos, unsigned endPos,
                         double diag_value = 0.0, double rhs_residual = 0.0) = 0;
};

class LinearSystem {
 public:
  virtual ~LinearSystem() = default;
  virtual void buildEdgeToNodeGraph() = 0;
  // nodes whose rows a DirichletBC algorithm owns
  // (LinearSystem::buildDirichletNodeGraph, include/LinearSystem.h:122-127):
  // nothing for Tpetra, skipped rows for hypre
  virtual void buildDirichletNodeGraph(const std::vector<int32_t>& nodes)
  {
    (void)nodes;
  }
  virtual void finalizeLinearSystem() = 0;
  virtual void zeroSystem() = 0;
  virtual CoeffApplier* get_coeff_applier() = 0;
  virtual void sumInto(unsigned numEntities, const int32_t* entities,
                       const double* rhs, const double* lhs) = 0;
  virtual void applyDirichletBCs(int solutionField, int bcValuesField,
                                 const std::vector<int32_t>& nodes,
                                 unsigned beginPos, unsigned endPos) = 0;
  virtual void resetRows(const std::vector<int32_t>& nodeList,
                         unsigned beginPos, unsigned endPos,
                         double diag_value = 0.0,
                         double rhs_residual = 0.0) = 0;
  virtual void loadComplete() = 0;
  virtual int solve(const nw_solver_cfg& cfg) = 0;
  int linearSolveIterations() const { return iters_; }
  double linearResidual() const { return linearResidual_; }
  double nonLinearResidual() const { return nonLinearResidual_; }
  unsigned numDof() const { return numDof_; }

 protected:
  unsigned numDof_ = 1;
  int iters_ = 0;
  double linearResidual_ = 0.0, nonLinearResidual_ = 0.0;
};

// The GPU-resident system of one Realm (TpetraLinearSystem analogue).
class GpuLinearSystem : public LinearSystem {
 public:
  GpuLinearSystem(Realm& realm, unsigned numDof);
  ~GpuLinearSystem() override;
  void buildEdgeToNodeGraph() override;
  void finalizeLinearSystem() override {}
  void zeroSystem() override;
  CoeffApplier* get_coeff_applier() override { return applier_.get(); }
  void sumInto(unsigned numEntities, const int32_t* entities,
               const double* rhs, const double* lhs) override;
  void applyDirichletBCs(int solutionField, int bcValuesField,
                         const std::vector<int32_t>& nodes, unsigned beginPos,
                         unsigned endPos) override;
  void resetRows(const std::vector<int32_t>& nodeList, unsigned beginPos,
                 unsigned endPos, double diag_value = 0.0,
                 double rhs_residual = 0.0) override;
  void loadComplete() override;
  int solve(const nw_solver_cfg& cfg) override;

  // Tpetra point-CSR view of the owned matrix and rhs (host copies).
  void download(std::vector<int64_t>& rowOffsets, std::vector<int32_t>& cols,
                std::vector<double>& vals, std::vector<double>& rhs) const;
  std::vector<double> solution() const;
  Realm& realm() { return realm_; }
  nw_solve_stats lastStats() const { return stats_; }
  bool isActive() const { return realm_.activeSystem() == this; }
  // throws unless the context holds this system's graph
  void requireActive(const char* what) const;

 protected:
  // the graph of this system into the context (both builds and switches)
  virtual void activate();
  Realm& realm_;
  std::unique_ptr<CoeffApplier> applier_;
  nw_solve_stats stats_{};
};

// HypreLinearSystem semantics (src/HypreLinearSystem.C) over the same
// device system: Dirichlet nodes are skipped rows that the assembly and
// sumInto leave alone (sum_into_1DoF, :2165-2239) until applyDirichletBCs
// writes their diagonal and rhs (:2408-2456), and the reported linear
// residual is hypre's relative residual times ||b|| (:2533; hypre returns
// the relative norm, src/HypreDirectSolver.C:124-130).
class HypreLinearSystem : public GpuLinearSystem {
 public:
  using GpuLinearSystem::GpuLinearSystem;
  void buildDirichletNodeGraph(const std::vector<int32_t>& nodes) override;
  void finalizeLinearSystem() override;
  int solve(const nw_solver_cfg& cfg) override;

 protected:
  void activate() override;

 private:
  std::vector<int32_t> skipped_;
};

// LinearSystem::create (src/LinearSystem.C:117-159): the backend by type
enum class LinearSystemType { Tpetra, Hypre };

class EquationSystem;

// SolverAlgorithm (include/SolverAlgorithm.h:30-101)
class SolverAlgorithm {
 public:
  SolverAlgorithm(Realm& realm, EquationSystem* eqSystem)
    : realm_(realm), eqSystem_(eqSystem)
  {
  }
  virtual ~SolverAlgorithm() = default;
  virtual void execute() = 0;
  virtual void initialize_connectivity() = 0;

 protected:
  Realm& realm_;
  EquationSystem* eqSystem_;
};

// AssembleEdgeSolverAlgorithm: the whole edge set in one launch.
class AssembleEdgeSolverAlgorithm : public SolverAlgorithm {
 public:
  using SolverAlgorithm::SolverAlgorithm;
  void initialize_connectivity() override;
};

class ScalarEdgeSolverAlg : public AssembleEdgeSolverAlgorithm {
 public:
  ScalarEdgeSolverAlg(Realm& realm, EquationSystem* eqSystem,
                      std::string dofName)
    : AssembleEdgeSolverAlgorithm(realm, eqSystem), dofName_(std::move(dofName))
  {
  }
  void execute() override;

 private:
  std::string dofName_;
};

class MomentumEdgeSolverAlg : public AssembleEdgeSolverAlgorithm {
 public:
  using AssembleEdgeSolverAlgorithm::AssembleEdgeSolverAlgorithm;
  void execute() override;
};

class ContinuityEdgeSolverAlg : public AssembleEdgeSolverAlgorithm {
 public:
  using AssembleEdgeSolverAlgorithm::AssembleEdgeSolverAlgorithm;
  void execute() override;
};

// MomentumEdgePecletAlg: writes the edge peclet_factor field
class MomentumEdgePecletAlg {
 public:
  explicit MomentumEdgePecletAlg(Realm& realm) : realm_(realm) {}
  void execute();

 private:
  Realm& realm_;
};

// GeometryInteriorAlg (src/ngp_algorithms/GeometryInteriorAlg.C:70-225):
// the edge area vectors and dual nodal volumes of the realm's elements
// (type [nelem]: 0 Hex8, 1 Tet4, 2 Wed6, 3 Pyr5, or null for all Hex8; conn
// [nelem][8]) from its coordinates, on the device.  The constructor hands
// the elements over once; execute() recomputes after the nodes move.
class GeometryInteriorAlg {
 public:
  GeometryInteriorAlg(Realm& realm, int64_t nelem, const int32_t* type,
                      const int32_t* conn);
  // the elements of the realm's mesh file (Realm::create_mesh)
  explicit GeometryInteriorAlg(Realm& realm);
  void execute();

 private:
  Realm& realm_;
};

// MdotEdgeAlg (src/ngp_algorithms/MdotEdgeAlg.C:43-194): writes the edge
// mass_flow_rate field from the new pressure (interior edges)
class MdotEdgeAlg {
 public:
  explicit MdotEdgeAlg(Realm& realm) : realm_(realm) {}
  void execute();

 private:
  Realm& realm_;
};

// NodalGradEdgeAlg (src/ngp_algorithms/NodalGradEdgeAlg.C:58-111): the
// interior-edge nodal gradient of node field `phi` into `gradPhi`
// (NW_F_PRESSURE -> NW_F_DPDX, NW_F_VELOCITY -> NW_F_DUDX, ...); needs the
// context's edge graph (built by the linear system).  It overwrites
// gradPhi unless `accumulate`: with boundary-face contributions added
// before it (the reference driver zeroes the field and every algorithm
// adds), construct it with accumulate = true.
class NodalGradEdgeAlg {
 public:
  NodalGradEdgeAlg(Realm& realm, int phi, int gradPhi, bool accumulate = false)
    : realm_(realm), phi_(phi), gradPhi_(gradPhi), accumulate_(accumulate)
  {
  }
  void execute();

 private:
  Realm& realm_;
  int phi_, gradPhi_;
  bool accumulate_;
};

// DirichletBC (src/DirichletBC.C): its nodes' rows become
// solution = bc value through applyDirichletBCs of the equation system's
// linear system; initialize_connectivity registers the nodes
// (buildDirichletNodeGraph, the hypre backend's skipped rows).
class DirichletBC : public SolverAlgorithm {
 public:
  DirichletBC(Realm& realm, EquationSystem* eqSystem, int field, int bcValues,
              std::vector<int32_t> nodes, unsigned beginPos, unsigned endPos)
    : SolverAlgorithm(realm, eqSystem), field_(field), bcValues_(bcValues),
      nodes_(std::move(nodes)), beginPos_(beginPos), endPos_(endPos)
  {
  }
  void execute() override;
  void initialize_connectivity() override;

 private:
  int field_, bcValues_;
  std::vector<int32_t> nodes_;
  unsigned beginPos_, endPos_;
};

class FixPressureAtNodeAlgorithm : public SolverAlgorithm {
 public:
  FixPressureAtNodeAlgorithm(Realm& realm, EquationSystem* eqSystem,
                             int32_t node, double refPressure)
    : SolverAlgorithm(realm, eqSystem), node_(node), refPressure_(refPressure)
  {
  }
  void execute() override;
  void initialize_connectivity() override {}

 private:
  int32_t node_;
  double refPressure_;
};

// EquationSystem::assemble_and_solve with the reference's phase timers.
class EquationSystem {
 public:
  EquationSystem(Realm& realm, unsigned numDof,
                 LinearSystemType type = LinearSystemType::Tpetra)
    : realm_(realm),
      linsys_(type == LinearSystemType::Hypre
                ? new HypreLinearSystem(realm, numDof)
                : new GpuLinearSystem(realm, numDof))
  {
  }
  void add_algorithm(std::unique_ptr<SolverAlgorithm> alg)
  {
    algs_.push_back(std::move(alg));
  }
  void initialize();
  // zeroSystem -> algorithms -> loadComplete -> solve (when cfg is given)
  void assemble_and_solve(const nw_solver_cfg* cfg);
  GpuLinearSystem& linsys() { return *linsys_; }
  double timerAssemble_ = 0, timerLoadComplete_ = 0, timerSolve_ = 0;

 private:
  Realm& realm_;
  std::unique_ptr<GpuLinearSystem> linsys_;
  std::vector<std::unique_ptr<SolverAlgorithm>> algs_;
};

}  // namespace nalu_amd
