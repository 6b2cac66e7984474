// nalu_adapter.cpp -- the reference-shaped host classes over the C ABI.
#include "nalu_adapter.h"

#include <algorithm>
#include <chrono>
#include <memory>
#include <string>

namespace nalu_amd {

void check(int rc, const char* what)
{
  if (rc != NW_OK)
    throw std::runtime_error(std::string(what) + ": " + nw_last_error());
}

nw_peclet SolutionOptions::peclet(const std::string& dof) const
{
  // EquationSystem::ngp_create_peclet_function (include/EquationSystem.h:
  // 401-417): classic form with the hybrid factor, or tanh(trans, width)
  nw_peclet p{};
  p.form = tanhFormDefault == "classic" ? 0 : 1;
  p.hybrid = get(hybridMap, dof, hybridDefault);
  p.c1 = tanhTransDefault;
  p.c2 = tanhWidthDefault;
  return p;
}

Realm::Realm(int device) { check(nw_ctx_create(device, &ctx_), "nw_ctx_create"); }

Realm::~Realm() { nw_ctx_destroy(ctx_); }

void Realm::set_mesh(int64_t nnode, const std::vector<int32_t>& edge_nodes)
{
  check(nw_mesh_set(ctx_, nnode, (int64_t)edge_nodes.size() / 2,
                    edge_nodes.data()),
        "nw_mesh_set");
  nnode_ = nnode;
  nedge_ = (int64_t)edge_nodes.size() / 2;
  active_ = nullptr;  // the new mesh has no graph yet
}

void Realm::create_mesh(const std::string& path)
{
  nw_exo* exo = nullptr;
  check(nw_exo_open(path.c_str(), &exo), "Realm::create_mesh");
  std::unique_ptr<nw_exo, void (*)(nw_exo*)> hold(exo, nw_exo_close);
  nw_exo_info info{};
  check(nw_exo_sizes(exo, &info), "Realm::create_mesh");
  std::vector<double> coords(3 * info.nnode);
  std::vector<int32_t> type(info.nelem), conn(8 * info.nelem);
  check(nw_exo_read_mesh(exo, coords.data(), type.data(), conn.data(), nullptr),
        "Realm::populate_mesh");
  int64_t ne = 0;
  check(nw_mixed_edges(info.nelem, type.data(), conn.data(), info.nnode, nullptr, 0, &ne),
        "Realm::create_edges");
  std::vector<int32_t> edges(2 * ne);
  check(nw_mixed_edges(info.nelem, type.data(), conn.data(), info.nnode, edges.data(), ne,
                       &ne),
        "Realm::create_edges");
  set_mesh(info.nnode, edges);
  set_field(NW_F_COORDINATES, coords);
  meshPath_ = path;
  elemType_ = std::move(type);
  elemConn_ = std::move(conn);
}

std::vector<int32_t> Realm::bc_nodes(const std::string& name) const
{
  if (meshPath_.empty()) throw std::runtime_error("Realm::bc_nodes: no mesh file");
  nw_exo* exo = nullptr;
  check(nw_exo_open(meshPath_.c_str(), &exo), "Realm::bc_nodes");
  std::unique_ptr<nw_exo, void (*)(nw_exo*)> hold(exo, nw_exo_close);
  int32_t idx = -1;
  int64_t count = 0;
  std::vector<int32_t> nodes;
  if (nw_exo_find(exo, NW_EXO_NODE_SET, name.c_str(), &idx) == NW_OK) {
    check(nw_exo_entity(exo, NW_EXO_NODE_SET, idx, nullptr, &count, nullptr, nullptr, 0),
          "Realm::bc_nodes");
    nodes.resize(count);
    check(nw_exo_node_set(exo, idx, nodes.data()), "Realm::bc_nodes");
    std::sort(nodes.begin(), nodes.end());
    nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
    return nodes;
  }
  check(nw_exo_find(exo, NW_EXO_SIDE_SET, name.c_str(), &idx), "Realm::bc_nodes");
  check(nw_exo_entity(exo, NW_EXO_SIDE_SET, idx, nullptr, &count, nullptr, nullptr, 0),
        "Realm::bc_nodes");
  std::vector<int32_t> elem(count), side(count);
  check(nw_exo_side_set(exo, idx, elem.data(), side.data()), "Realm::bc_nodes");
  int64_t nn = 0;
  check(nw_sides_to_nodes(nelem(), elemType_.data(), elemConn_.data(), nnode_, count,
                          elem.data(), side.data(), nullptr, &nn),
        "Realm::bc_nodes");
  nodes.resize(nn);
  check(nw_sides_to_nodes(nelem(), elemType_.data(), elemConn_.data(), nnode_, count,
                          elem.data(), side.data(), nodes.data(), &nn),
        "Realm::bc_nodes");
  return nodes;
}

void Realm::write_output(const std::string& path,
                         const std::vector<std::pair<std::string, int>>& fields,
                         double time) const
{
  std::vector<std::string> names;
  std::vector<double> values;
  for (const auto& f : fields) {
    const int id = f.second, nc = nw_field_components(id);
    if (nc <= 0 || id == NW_E_AREA_VECTOR || id == NW_E_MASS_FLOW_RATE ||
        id == NW_E_PECLET_FACTOR || id == NW_E_PECLET_NUMBER)
      throw std::runtime_error("Realm::write_output: " + f.first + " is not a node field");
    const std::vector<double> v = get_field(id);
    for (int c = 0; c < nc; ++c) {
      std::string name = f.first;
      if (nc == 3) name += std::string("_") + "xyz"[c];
      if (nc == 9) name += std::string("_") + "xyz"[c / 3] + "xyz"[c % 3];
      names.push_back(name);
      for (int64_t n = 0; n < nnode_; ++n) values.push_back(v[n * nc + c]);
    }
  }
  std::vector<const char*> cn;
  for (const auto& n : names) cn.push_back(n.c_str());
  const std::vector<double> coords = get_field(NW_F_COORDINATES);
  check(nw_exo_write(path.c_str(), "nalu-wind edge realm", nnode_, coords.data(), nelem(),
                     elemType_.data(), elemConn_.data(), nullptr, nullptr,
                     (int32_t)names.size(), cn.data(), values.data(), time, nullptr),
        "Realm::write_output");
}

void Realm::set_field(int field, const std::vector<double>& data)
{
  check(nw_field_upload(ctx_, field, data.data(), (int64_t)data.size()),
        "nw_field_upload");
}

std::vector<double> Realm::get_field(int field) const
{
  const bool edge = field == NW_E_AREA_VECTOR || field == NW_E_MASS_FLOW_RATE ||
                    field == NW_E_PECLET_FACTOR || field == NW_E_PECLET_NUMBER;
  std::vector<double> out((edge ? nedge_ : nnode_) * nw_field_components(field));
  check(nw_field_download(ctx_, field, out.data(), (int64_t)out.size()),
        "nw_field_download");
  return out;
}

namespace {

class GpuCoeffApplier : public CoeffApplier {
 public:
  explicit GpuCoeffApplier(nw_ctx* ctx) : ctx_(ctx) {}
  void operator()(unsigned numEntities, const int32_t* entities,
                  const double* rhs, const double* lhs) override
  {
    check(nw_linsys_sum_into(ctx_, 1, (int)numEntities, entities, lhs, rhs),
          "sumInto");
  }
  void resetRows(unsigned numNodes, const int32_t* nodeList, unsigned beginPos,
                 unsigned endPos, double diag_value,
                 double rhs_residual) override
  {
    check(nw_linsys_reset_rows(ctx_, numNodes, nodeList, (int)beginPos,
                               (int)endPos, diag_value, rhs_residual),
          "resetRows");
  }

 private:
  nw_ctx* ctx_;
};

double now()
{
  return std::chrono::duration<double>(
           std::chrono::steady_clock::now().time_since_epoch())
    .count();
}

}  // namespace

GpuLinearSystem::GpuLinearSystem(Realm& realm, unsigned numDof)
  : realm_(realm), applier_(new GpuCoeffApplier(realm.ctx()))
{
  numDof_ = numDof;
}

GpuLinearSystem::~GpuLinearSystem()
{
  if (isActive()) realm_.setActiveSystem(nullptr);
}

void GpuLinearSystem::requireActive(const char* what) const
{
  if (!isActive())
    throw std::runtime_error(
      std::string(what) +
      ": another linear system of this Realm holds the context (call "
      "zeroSystem() of this system first)");
}

void GpuLinearSystem::activate()
{
  check(nw_linsys_build_edge_graph(realm_.ctx(), (int)numDof_),
        "buildEdgeToNodeGraph");
  realm_.setActiveSystem(this);
}

void GpuLinearSystem::buildEdgeToNodeGraph() { activate(); }

void GpuLinearSystem::zeroSystem()
{
  if (!isActive()) activate();
  check(nw_linsys_zero(realm_.ctx()), "zeroSystem");
}

void GpuLinearSystem::sumInto(unsigned numEntities, const int32_t* entities,
                              const double* rhs, const double* lhs)
{
  requireActive("sumInto");
  (*applier_)(numEntities, entities, rhs, lhs);
}

void GpuLinearSystem::applyDirichletBCs(int solutionField, int bcValuesField,
                                        const std::vector<int32_t>& nodes,
                                        unsigned beginPos, unsigned endPos)
{
  requireActive("applyDirichletBCs");
  check(nw_linsys_apply_dirichlet(realm_.ctx(), (int64_t)nodes.size(),
                                  nodes.data(), solutionField, bcValuesField,
                                  (int)beginPos, (int)endPos),
        "applyDirichletBCs");
}

void GpuLinearSystem::resetRows(const std::vector<int32_t>& nodeList,
                                unsigned beginPos, unsigned endPos,
                                double diag_value, double rhs_residual)
{
  requireActive("resetRows");
  applier_->resetRows((unsigned)nodeList.size(), nodeList.data(), beginPos,
                      endPos, diag_value, rhs_residual);
}

void GpuLinearSystem::loadComplete()
{
  requireActive("loadComplete");
  check(nw_linsys_load_complete(realm_.ctx()), "loadComplete");
}

int GpuLinearSystem::solve(const nw_solver_cfg& cfg)
{
  requireActive("solve");
  const int rc = nw_linsys_solve(realm_.ctx(), &cfg, &stats_);
  if (rc != NW_OK && rc != NW_E_NOCONV) check(rc, "solve");
  iters_ = stats_.iters;
  linearResidual_ = stats_.final_abs;  // ||A x - b||_2 (LinearSolver.C:251)
  nonLinearResidual_ = stats_.bnorm;   // ||rhs||_2 (TpetraLinearSystem.C:1853-1859)
  return rc;
}

void GpuLinearSystem::download(std::vector<int64_t>& rowOffsets,
                               std::vector<int32_t>& cols,
                               std::vector<double>& vals,
                               std::vector<double>& rhs) const
{
  requireActive("download");
  int64_t nrows = 0, nnz = 0;
  check(nw_linsys_sizes(realm_.ctx(), &nrows, &nnz), "sizes");
  rowOffsets.resize(nrows + 1);
  cols.resize(nnz);
  vals.resize(nnz);
  rhs.resize(nrows);
  check(nw_linsys_download(realm_.ctx(), rowOffsets.data(), cols.data(),
                           vals.data(), rhs.data()),
        "download");
}

std::vector<double> GpuLinearSystem::solution() const
{
  requireActive("solution");
  int64_t nrows = 0;
  check(nw_linsys_sizes(realm_.ctx(), &nrows, nullptr), "sizes");
  std::vector<double> x(nrows);
  check(nw_linsys_solution_download(realm_.ctx(), x.data()), "solution");
  return x;
}

void HypreLinearSystem::buildDirichletNodeGraph(const std::vector<int32_t>& nodes)
{
  skipped_.insert(skipped_.end(), nodes.begin(), nodes.end());
}

void HypreLinearSystem::activate()
{
  GpuLinearSystem::activate();
  check(nw_linsys_set_skipped_rows(realm_.ctx(), (int64_t)skipped_.size(),
                                   skipped_.data()),
        "buildDirichletNodeGraph");
}

void HypreLinearSystem::finalizeLinearSystem()
{
  // the skipped rows gathered by initialize_connectivity
  requireActive("finalizeLinearSystem");
  check(nw_linsys_set_skipped_rows(realm_.ctx(), (int64_t)skipped_.size(),
                                   skipped_.data()),
        "buildDirichletNodeGraph");
}

int HypreLinearSystem::solve(const nw_solver_cfg& cfg)
{
  const int rc = GpuLinearSystem::solve(cfg);
  // hypre reports the relative residual; the linear system scales it by
  // ||b|| to compare with the Tpetra report (HypreLinearSystem.C:2525-2533)
  linearResidual_ = stats_.final_rel * stats_.bnorm;
  return rc;
}

namespace {
// throws unless the equation system's linear system holds the context
void require_system(EquationSystem* eq, const char* what)
{
  if (eq) eq->linsys().requireActive(what);
}
}  // namespace

void AssembleEdgeSolverAlgorithm::initialize_connectivity()
{
  // the edge graph is built once by the linear system of the equation
  // system (buildEdgeToNodeGraph); nothing per algorithm
}

void ScalarEdgeSolverAlg::execute()
{
  const SolutionOptions& so = realm_.solutionOptions;
  nw_scalar_opts o{};
  o.alpha = so.get(so.alphaMap, dofName_, so.alphaDefault);
  o.alpha_upw = so.get(so.alphaUpwMap, dofName_, so.alphaUpwDefault);
  o.upw = so.get(so.upwMap, dofName_, so.upwDefault);
  o.relax = so.get(so.relaxMap, dofName_, so.relaxDefault);
  auto it = so.limiterMap.find(dofName_);
  o.use_limiter = (it != so.limiterMap.end() && it->second) ? 1 : 0;
  o.peclet = so.peclet(dofName_);
  o.accumulate = 1;  // sumInto semantics after zeroSystem
  require_system(eqSystem_, "ScalarEdgeSolverAlg");
  check(nw_assemble_scalar(realm_.ctx(), &o), "ScalarEdgeSolverAlg");
}

void MomentumEdgeSolverAlg::execute()
{
  const SolutionOptions& so = realm_.solutionOptions;
  const std::string dof = "velocity";
  nw_momentum_opts o{};
  o.alpha = so.get(so.alphaMap, dof, so.alphaDefault);
  o.alpha_upw = so.get(so.alphaUpwMap, dof, so.alphaUpwDefault);
  o.upw = so.get(so.upwMap, dof, so.upwDefault);
  o.relax = so.get(so.relaxMap, dof, so.relaxDefault);
  o.include_divU = so.includeDivU;
  auto it = so.limiterMap.find(dof);
  o.use_limiter = (it != so.limiterMap.end() && it->second) ? 1 : 0;
  o.accumulate = 1;
  require_system(eqSystem_, "MomentumEdgeSolverAlg");
  check(nw_assemble_momentum(realm_.ctx(), &o), "MomentumEdgeSolverAlg");
}

void ContinuityEdgeSolverAlg::execute()
{
  const SolutionOptions& so = realm_.solutionOptions;
  nw_continuity_opts o{};
  o.dt = realm_.timeIntegrator.timeStepN;
  o.gamma1 = realm_.timeIntegrator.gamma1;
  auto it = so.nocMap.find("pressure");
  o.noc = (it == so.nocMap.end() ? so.nocDefault : it->second) ? 1.0 : 0.0;
  o.interp_together = so.mdotInterpRhoUTogether ? 1.0 : 0.0;
  o.incompressible = so.solveIncompressibleContinuity ? 1.0 : 0.0;
  o.accumulate = 1;
  o.balanced_buoyancy = so.use_balanced_buoyancy_force_ ? 1 : 0;
  for (int d = 0; d < 3; ++d) o.gravity[d] = so.gravity[d];
  require_system(eqSystem_, "ContinuityEdgeSolverAlg");
  check(nw_assemble_continuity(realm_.ctx(), &o), "ContinuityEdgeSolverAlg");
}

void MomentumEdgePecletAlg::execute()
{
  const nw_peclet p = realm_.solutionOptions.peclet("velocity");
  check(nw_momentum_edge_peclet(realm_.ctx(), &p), "MomentumEdgePecletAlg");
}

GeometryInteriorAlg::GeometryInteriorAlg(Realm& realm, int64_t nelem,
                                         const int32_t* type, const int32_t* conn)
  : realm_(realm)
{
  check(nw_mesh_elements(realm_.ctx(), nelem, type, conn), "GeometryInteriorAlg");
}

GeometryInteriorAlg::GeometryInteriorAlg(Realm& realm)
  : GeometryInteriorAlg(realm, realm.nelem(), realm.elemTypes().data(),
                        realm.elemConn().data())
{
}

void GeometryInteriorAlg::execute()
{
  check(nw_compute_geometry(realm_.ctx()), "GeometryInteriorAlg");
}

void MdotEdgeAlg::execute()
{
  const SolutionOptions& so = realm_.solutionOptions;
  auto it = so.nocMap.find("pressure");
  const bool noc = it == so.nocMap.end() ? so.nocDefault : it->second;
  nw_mdot_opts o{};
  o.noc = noc ? 1.0 : 0.0;
  o.interp_together = so.mdotInterpRhoUTogether ? 1.0 : 0.0;
  o.balanced_buoyancy = so.use_balanced_buoyancy_force_ ? 1 : 0;
  for (int d = 0; d < 3; ++d) o.gravity[d] = so.gravity[d];
  check(nw_compute_mdot(realm_.ctx(), &o), "MdotEdgeAlg");
}

void NodalGradEdgeAlg::execute()
{
  check(nw_nodal_grad(realm_.ctx(), phi_, gradPhi_, accumulate_ ? 1 : 0),
        "NodalGradEdgeAlg");
}

void DirichletBC::initialize_connectivity()
{
  if (eqSystem_) eqSystem_->linsys().buildDirichletNodeGraph(nodes_);
}

void DirichletBC::execute()
{
  if (!eqSystem_) throw std::runtime_error("DirichletBC without an equation system");
  eqSystem_->linsys().applyDirichletBCs(field_, bcValues_, nodes_, beginPos_,
                                        endPos_);
}

void FixPressureAtNodeAlgorithm::execute()
{
  require_system(eqSystem_, "FixPressureAtNodeAlgorithm");
  check(nw_fix_pressure_at_node(realm_.ctx(), node_, refPressure_),
        "FixPressureAtNodeAlgorithm");
}

void EquationSystem::initialize()
{
  linsys_->buildEdgeToNodeGraph();
  for (auto& a : algs_) a->initialize_connectivity();
  linsys_->finalizeLinearSystem();
}

void EquationSystem::assemble_and_solve(const nw_solver_cfg* cfg)
{
  // src/EquationSystem.C:301-350: zero, all algorithms, loadComplete, solve
  double t0 = now();
  linsys_->zeroSystem();
  for (auto& a : algs_) a->execute();
  check(nw_ctx_synchronize(realm_.ctx()), "synchronize");
  double t1 = now();
  timerAssemble_ += t1 - t0;
  linsys_->loadComplete();
  check(nw_ctx_synchronize(realm_.ctx()), "synchronize");
  double t2 = now();
  timerLoadComplete_ += t2 - t1;
  if (cfg) {
    linsys_->solve(*cfg);
    timerSolve_ += now() - t2;
  }
}

}  // namespace nalu_amd
