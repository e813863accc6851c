// Host C++ API of the reference's classes, forwarding to the C-ABI.
#include "epic_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

namespace epic2d {

Rnd rnd;

bool string_to_bool(const std::string& s) {
  return s == "true" || s == "True" || s == "T" || s == "t" || s == "1" || s == "yes" || s == "Yes" || s == "Y" ||
         s == "y";
}

// ---- INIParser (iniparser.cpp:4-113) ----------------------------------------
INIParser::Sections INIParser::parse_text(const std::string& text) {
  Sections sections;
  std::string currentSection;
  std::istringstream file(text);
  std::string line;
  while (std::getline(file, line)) {
    if (line.empty() || line[0] == ';' || line[0] == '#') continue;
    if (line[0] == '[' && line[line.length() - 1] == ']') {
      currentSection = line.substr(1, line.length() - 2);
      continue;
    }
    if (currentSection == "species") {
      sections[currentSection][std::to_string(sections[currentSection].size())] = line;
    } else {
      std::istringstream iss(line);
      std::string key, value;
      if (std::getline(iss, key, '=') && std::getline(iss, value)) {
        trim(key);
        trim(value);
        sections[currentSection][key] = value;
      }
    }
  }
  return sections;
}

INIParser::Sections INIParser::parse(const std::string& filename) {
  std::ifstream f(filename);
  if (!f.is_open()) throw std::runtime_error("Error opening file: " + filename);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse_text(ss.str());
}

int INIParser::getInt(const std::map<std::string, std::string>& section, const std::string& key) {
  return std::stoi(section.at(key));
}
double INIParser::getDouble(const std::map<std::string, std::string>& section, const std::string& key) {
  return std::stod(section.at(key));
}
std::string INIParser::getString(const std::map<std::string, std::string>& section, const std::string& key) {
  return section.at(key);
}
void INIParser::trim(std::string& str) {
  if (str.empty()) return;
  const size_t first = str.find_first_not_of(" \t");
  const size_t last = str.find_last_not_of(" \t");
  if (first == std::string::npos) str.clear();
  else str = str.substr(first, last - first + 1);
}
std::vector<std::string> INIParser::split(const std::string& str, char delimiter) {
  std::vector<std::string> tokens;
  std::string token;
  std::istringstream s(str);
  while (std::getline(s, token, delimiter)) tokens.push_back(token);
  return tokens;
}
std::pair<std::string, int> INIParser::loadtypeextract(const std::string& position_init) {
  std::stringstream ss(position_init);
  std::string func_type, n_str;
  if (std::getline(ss, func_type, '(') && std::getline(ss, n_str, ')')) return {func_type, std::stoi(n_str)};
  return {position_init, 0};
}

// ---- Domain -------------------------------------------------------------------
Domain::Domain(double x0, double y0, double dx, double dy, int nx, int ny)
    : x0(x0), y0(y0), dx(dx), dy(dy), nx(nx), ny(ny) {
  Lx = (nx - 1) * dx;  // domain.cpp:14-15
  Ly = (ny - 1) * dy;
}

Domain::~Domain() {
  if (ctx_) epic_destroy(ctx_);
}

void Domain::set_normparam(double LDe_, double wpe_, double LDi_, double wpi_) {
  LDe = LDe_;
  wpe = wpe_;
  LDi = LDi_;
  wpi = wpi_;
}

// domain.cpp:123-183.  vel_normscheme is whatever the caller set; the
// reference's main never sets it, and its trailing else gives L*W.
void Domain::set_normscheme() {
  if (normscheme == 1 || normscheme == 3) {
    L = LDe;
    W = wpe;
  } else if (normscheme == 2 || normscheme == 4) {
    L = LDi;
    W = (normscheme == 2) ? wpi : wpe;
  } else if (normscheme == 5) {
    LDe = lenght_scale;
    wpe = time_scale;
    L = LDe;
    W = wpe;
  } else {
    throw std::invalid_argument("Invalid normscheme flag");
  }
  vel_norm = (vel_normscheme == 3) ? IAW_vel : L * W;
}

void Domain::set_time(double DT_, int NUM_TS_, int wi, int wip) {
  DT = DT_;
  NUM_TS = NUM_TS_;
  write_interval = wi;
  write_interval_phase = wip;
}

int Domain::register_species(Species* s) {
  reg_.push_back(s);
  return (int)reg_.size() - 1;
}

void Domain::bind_species(std::vector<Species>& list) {
  reg_.clear();
  for (auto& s : list) reg_.push_back(&s);
}

void Domain::set_grids(const std::vector<Grid*>& grids) {
  if (grids_set_ && grids == grids_) return;
  std::vector<epic_conductor_desc> d;
  for (Grid* g : grids) d.push_back(g->desc());
  check(epic_set_conductors(ctx(), d.data(), (int)d.size()), "epic_set_conductors");
  grids_ = grids;
  grids_set_ = true;
}

void Domain::check(int rc, const char* what) {
  if (rc != EPIC_OK) {
    std::string m = std::string(what) + ": " + (ctx_ ? epic_last_error(ctx_) : "no context");
    throw EpicError(rc, m);
  }
}

epic_ctx* Domain::ctx() {
  if (ctx_) return ctx_;
  if (reg_.empty()) throw EpicError(EPIC_ERR_STATE, "no species registered with the domain");
  epic_domain_desc d{};
  d.x0 = x0; d.y0 = y0; d.dx = dx; d.dy = dy; d.nx = nx; d.ny = ny;
  d.L = L; d.W = W; d.LDe = LDe; d.wpe = wpe; d.vel_norm = vel_norm; d.density = density; d.DT = DT;
  if (bc == "pbc") d.bc = EPIC_BC_PBC;
  else if (bc == "open") d.bc = EPIC_BC_OPEN;
  else if (bc == "reflective") d.bc = EPIC_BC_REFLECTIVE;
  else throw EpicError(EPIC_ERR_ARG, "unknown bc " + bc);
  if (shapefunction == "CIC") d.shape = EPIC_SHAPE_CIC;
  else if (shapefunction == "NGP") d.shape = EPIC_SHAPE_NGP;
  else throw EpicError(EPIC_ERR_ARG, "unknown shapefunction " + shapefunction);
  if (SolverType == "gs") d.solver = EPIC_SOLVER_GS;
  else if (SolverType == "gshybrid") d.solver = EPIC_SOLVER_GS_BOLTZMANN;
  else if (SolverType == "hybrid") d.solver = EPIC_SOLVER_NEWTON_CG;
  else if (SolverType == "cg" || SolverType == "iter" || SolverType == "iterative") d.solver = EPIC_SOLVER_PCG;
  else if (SolverType == "pcg_mixed") d.solver = EPIC_SOLVER_PCG_MIXED;
  else if (SolverType == "spectral") d.solver = EPIC_SOLVER_SPECTRAL;
  else throw EpicError(EPIC_ERR_ARG, "solver " + SolverType + " is not part of this build");
  d.poisson_op = poisson_op >= 0 ? poisson_op
                                 : (d.solver == EPIC_SOLVER_GS || d.solver == EPIC_SOLVER_GS_BOLTZMANN
                                        ? EPIC_OP_DIRICHLET_WALL
                                        : EPIC_OP_REF_GHOST);
  d.tolerance = tolerance;
  d.tol_relative = tol_relative;
  d.max_iteration = max_iteration;
  d.sor_omega = 1.4;
  d.precision = precision;
  d.sort_interval = sort_interval;
  d.seed = seed;
  d.precond = precond;
  std::vector<epic_species_desc> sp(reg_.size());
  for (size_t k = 0; k < reg_.size(); k++) {
    const Species& s = *reg_[k];
    epic_species_desc& e = sp[k];
    std::memset(&e, 0, sizeof(e));
    std::strncpy(e.name, s.name.c_str(), sizeof(e.name) - 1);
    e.mass = s.mass; e.charge = s.charge; e.spwt = s.spwt; e.temp = s.temp;
    e.numparticle = s.numparticle;
    e.capacity = (int64_t)std::ceil(s.numparticle * capacity_factor) + 1;
    e.vsx = s.vsx; e.vsy = s.vsy; e.fract_den = s.fract_den;
    e.pushed = (s.name == "ion" && ionfixed == 1) ? 0 : 1;  // main.cpp:419
  }
  const int rc = epic_create(&d, sp.data(), (int)sp.size(), &ctx_);
  if (rc != EPIC_OK) throw EpicError(rc, "epic_create failed");
  for (auto& e : pending_emitters_) check(epic_add_emitter(ctx_, &e), "epic_add_emitter");
  pending_emitters_.clear();
  return ctx_;
}

std::vector<double> Domain::field(int which) {
  std::vector<double> out((size_t)nx * ny);
  check(epic_get_field(ctx(), which, out.data()), "epic_get_field");
  return out;
}

// Domain::ComputeRho (domain.cpp:194-202)
void Domain::ComputeRho(std::vector<Species>& species) {
  (void)species;
  check(epic_compute_rho(ctx()), "epic_compute_rho");
}

// Domain::Compute_PE (domain.cpp:266-286)
double Domain::Compute_PE(Species& species) {
  double pe = 0;
  check(epic_compute_pe(ctx(), species.index, &pe), "epic_compute_pe");
  return pe;
}

// ---- Species ------------------------------------------------------------------
Species::Species(std::string name, double mass, double charge, double spwt, double temp, int numparticle,
                 double vsx, double vsy, double fract_den, std::string initialization, Domain& domain)
    : name(std::move(name)), mass(mass), charge(charge), spwt(spwt), temp(temp), numparticle(numparticle),
      initialization(std::move(initialization)), vsx(vsx), vsy(vsy), fract_den(fract_den), domain_(&domain) {
  index = domain.register_species(this);
}

void Species::Push_species_serial(int sub_cycle, const std::vector<Grid*>& grids) {
  domain_->set_grids(grids);
  domain_->check(epic_push(domain_->ctx(), index, sub_cycle), "epic_push");
}
void Species::ScatterSpecies_serial() { domain_->check(epic_scatter(domain_->ctx(), index), "epic_scatter"); }
void Species::Rewind_species() { domain_->check(epic_rewind(domain_->ctx(), index), "epic_rewind"); }
std::array<double, 3> Species::Compute_KE(Species& species) {
  std::array<double, 3> ke{};
  domain_->check(epic_compute_ke(domain_->ctx(), index, species.index, ke.data()), "epic_compute_ke");
  return ke;
}
std::array<double, 3> Species::Compute_Momentum(Species& species) {
  std::array<double, 3> p{};
  domain_->check(epic_compute_momentum(domain_->ctx(), index, species.index, p.data()), "epic_compute_momentum");
  return p;
}
int64_t Species::count() {
  int64_t n = 0;
  domain_->check(epic_count(domain_->ctx(), index, &n), "epic_count");
  return n;
}
std::vector<double> Species::den() { return domain_->field(10 + index); }

void Species::particles(std::vector<double>& pos, std::vector<double>& vel) {
  epic_ctx* c = domain_->ctx();
  int64_t n = 0;
  domain_->check(epic_count(c, index, &n), "epic_count");
  std::vector<double> x(n), y(n), vx(n), vy(n), vz(n);
  std::vector<uint32_t> id(n);
  int64_t got = 0;
  domain_->check(epic_download_particles(c, index, x.data(), y.data(), vx.data(), vy.data(), vz.data(), id.data(), n,
                                         &got),
                 "epic_download_particles");
  pos.resize(2 * got);
  vel.resize(2 * got);
  for (int64_t k = 0; k < got; k++) {
    pos[2 * k] = x[k];
    pos[2 * k + 1] = y[k];
    vel[2 * k] = vx[k];
    vel[2 * k + 1] = vy[k];
  }
}

// ---- Init (init.cpp:3-114) -------------------------------------------------------
double Init::SampleVel(Species& species) { return SampleVel(species, species.temp); }
double Init::SampleVel(Species& species, double temp) {
  const double v_th = std::sqrt(2 * Const::K_b * temp * Const::EV_to_K / species.mass);
  const double r1 = rnd();  // draws sequenced left to right
  const double r2 = rnd();
  const double r3 = rnd();
  return v_th * std::sqrt(2) * (r1 + r2 + r3 - 1.5);
}

Init::Init(Species& species, Domain& domain, std::vector<Grid*>& grids) {
  auto lt = INIParser::loadtypeextract(species.initialization);
  const std::string& type = lt.first;
  const int n = lt.second;
  const int N = species.numparticle;
  std::vector<double> x(N), y(N), vx(N), vy(N), vz(N, 0.0);
  std::random_device rd;
  std::mt19937 gen(rd());
  std::normal_distribution<> gx((domain.x0 + domain.Lx) / 2.0, domain.Lx / 5);
  std::normal_distribution<> gy((domain.y0 + domain.Ly) / 2.0, domain.Ly / 5);
  // a position on a conductor is drawn again (init.cpp:71-96)
  auto on_grid = [&](double qx, double qy) {
    for (Grid* g : grids)
      if (g->IsGrid(qx, qy)) return true;
    return false;
  };
  for (int p = 0; p < N; p++) {
    double px = 0, py = 0;
  draw:
    if (type == "random") {
      px = domain.x0 + domain.Lx * rnd();
      py = domain.y0 + domain.Ly * rnd();
    } else if (type == "uniform") {
      const int nxp = (int)std::sqrt((double)N);
      px = domain.x0 + (p % nxp) * (domain.Lx / (nxp - 1));
      py = domain.y0 + (p / nxp) * (domain.Ly / (nxp - 1));
    } else if (type == "sin" || type == "cos") {
      px = domain.x0 + p * (domain.Lx / (N - 1));
      py = domain.y0 + p * (domain.Ly / (N - 1));
      const double kx = 2 * Const::PI * n / domain.Lx, ky = 2 * Const::PI * n / domain.Ly;
      if (type == "sin") { px += std::sin(kx * px); py += std::sin(ky * py); }
      else { px += std::cos(kx * px); py += std::cos(ky * py); }
    } else if (type == "gaussian") {
      px = gx(gen);
      py = gy(gen);
    } else {
      throw EpicError(EPIC_ERR_ARG, "unknown load type " + type);
    }
    if (on_grid(px, py)) {
      // the reference redraws a fixed position for ever
      if (type != "random" && type != "gaussian")
        throw EpicError(EPIC_ERR_ARG, "load type " + type + " places a particle on a conductor");
      goto draw;
    }
    double pvx = SampleVel(species) + species.vsx * domain.vel_norm;
    double pvy = SampleVel(species) + species.vsy * domain.vel_norm;
    x[p] = px;
    y[p] = py;
    vx[p] = pvx / domain.vel_norm;
    vy[p] = pvy / domain.vel_norm;
  }
  domain.check(epic_upload_particles(domain.ctx(), species.index, x.data(), y.data(), vx.data(), vy.data(),
                                     vz.data(), N),
               "epic_upload_particles");
}

// ---- internal conductors ---------------------------------------------------------
CircularGrid::CircularGrid(Domain& domain, int center_x, int center_y, int num_electrodes, double voltage,
                           double electrode_rad, double grid_rad)
    : Grid(domain), center_x(center_x), center_y(center_y), num_electrodes(num_electrodes), voltage(voltage),
      electrode_rad(electrode_rad), grid_rad(grid_rad) {}

// Electrode::Iselectrode (electrode.cpp:19-26) over the electrodes that exist
bool CircularGrid::IsGrid(double x, double y) {
  for (size_t k = 0; k < ex_.size(); k++) {
    const double dx = x - ex_[k], dy = y - ey_[k];
    if (dx * dx + dy * dy <= electrode_rad * electrode_rad) return true;
  }
  return false;
}

// grid.cpp:40-50 (the device writes the electrodes' voltages)
void CircularGrid::initialize() {
  ex_.clear();
  ey_.clear();
  for (int k = 0; k < num_electrodes; ++k) {
    const double theta = k * (2.0 * M_PI / num_electrodes);
    ex_.push_back(center_x + grid_rad * std::cos(theta));
    ey_.push_back(center_y + grid_rad * std::sin(theta));
  }
}

epic_conductor_desc CircularGrid::desc() const {
  epic_conductor_desc d{};
  d.kind = EPIC_COND_CIRCULAR;
  d.x_center = center_x;
  d.y_center = center_y;
  d.electrode_number = num_electrodes;
  d.grid_radius = grid_rad;
  d.electrode_radius = electrode_rad;
  d.voltage = voltage;
  return d;
}

ReactConductorGrid::ReactConductorGrid(int min_x, int max_x, int min_y, int max_y, double voltage, Domain& domain)
    : Grid(domain), min_x(min_x), max_x(max_x), min_y(min_y), max_y(max_y), voltage(voltage) {}

bool ReactConductorGrid::IsGrid(double x, double y) {
  return (x >= min_x && x <= max_x && y >= min_y && y <= max_y);
}

epic_conductor_desc ReactConductorGrid::desc() const {
  epic_conductor_desc d{};
  d.kind = EPIC_COND_RECT;
  d.min_x = min_x;
  d.max_x = max_x;
  d.min_y = min_y;
  d.max_y = max_y;
  d.voltage = voltage;
  return d;
}

Grid* Create_Grid(GridType type, Domain& domain, const GridParams& params, const RectangularGridParams& rparam) {
  switch (type) {
    case GridType::Circular:
      return new CircularGrid(domain, params.grid_centerx, params.grid_centery, params.num_electrodes,
                              params.voltage, params.electrode_rad, params.grid_rad);
    case GridType::Line:
    case GridType::ReactConductor:
      return new ReactConductorGrid(rparam.min_x, rparam.max_x, rparam.min_y, rparam.max_y, rparam.voltage, domain);
    default:
      throw EpicError(EPIC_ERR_ARG, "custom grids (electrodes added one by one) are not supported");
  }
}

// ---- FieldSolve -----------------------------------------------------------------
FieldSolve::FieldSolve(Domain& domain, const std::vector<Grid*>& grids) : domain(domain), grids(grids) {
  domain.set_grids(grids);
}

void FieldSolve::PotentialSolver() {
  int it = 0;
  double err = 0;
  // gssolver calls each grid's initialize() (field.cpp:100-103)
  if (domain.SolverType == "gs")
    for (Grid* g : grids) g->initialize();
  domain.check(epic_solve_potential(domain.ctx(), &it, &err), "epic_solve_potential");
  domain.error = err;
  domain.last_iterations = it;
}
void FieldSolve::CalculateEfield() { domain.check(epic_efield(domain.ctx()), "epic_efield"); }

// ---- cross sections (cross_section.cpp:4-39) ------------------------------------------
double compute_elastic_CS(double energy) {
  return (1e-20 * std::fabs(6.0 / std::pow(1.0 + (energy / 0.1) + std::pow(energy / 0.6, 2.0), 3.3) -
                            1.1 * std::pow(energy, 1.4) / (1.0 + std::pow(energy / 15.0, 1.2)) /
                                std::sqrt(1.0 + std::pow(energy / 5.5, 2.5) + std::pow(energy / 60.0, 4.1)) +
                            0.05 / std::pow(1.0 + energy / 10.0, 2.0) +
                            0.01 * std::pow(energy, 3.0) / (1.0 + std::pow(energy / 12.0, 6.0))));
}
double compute_excitation_CS(double energy) {
  const double th = 11.5;
  if (energy > th)
    return (1e-20 * (0.034 * std::pow(energy - th, 1.1) * (1.0 + std::pow(energy / 15.0, 2.8)) /
                         (1.0 + std::pow(energy / 23.0, 5.5)) +
                     0.023 * (energy - th) / std::pow(1.0 + energy / 80.0, 1.9)));
  return 0.0;
}
double compute_ionization_CS(double energy) {
  const double th = 15.8;
  if (energy > th)
    return (1e-20 * (970.0 * (energy - th) / std::pow(70.0 + energy, 2.0) +
                     0.06 * std::pow(energy - th, 2.0) * std::exp(-energy / 9)));
  return 0.0;
}

// ---- CollisionHandler ---------------------------------------------------------------
// The reference tabulates 5,000,000 linear points (collision.cpp:21-42); the
// device interpolates a log-spaced table of the same formulas.
static const int NTAB = 4096;
static const double E_MIN = 0.005, E_MAX = 25000.0;

void CollisionHandler::set_electron_cross_sections() {
  table_.assign((size_t)EPIC_MCC_NPROC * NTAB, 0.0);
  const double l0 = std::log(E_MIN), l1 = std::log(E_MAX);
  for (int k = 0; k < NTAB; k++) {
    const double e = std::exp(l0 + (l1 - l0) * k / (NTAB - 1));
    table_[EPIC_MCC_E_ELASTIC * NTAB + k] = domain.enable_elastic_collision ? compute_elastic_CS(e) : 0.0;
    table_[EPIC_MCC_E_EXCITATION * NTAB + k] = domain.enable_excitation_collision ? compute_excitation_CS(e) : 0.0;
    table_[EPIC_MCC_E_IONIZATION * NTAB + k] = domain.enable_ionization_collision ? compute_ionization_CS(e) : 0.0;
  }
}

void CollisionHandler::calc_total_cross_sections() {
  if (table_.empty()) set_electron_cross_sections();
  const bool any = domain.GAS_DENSITY > 0 &&
                   (domain.enable_elastic_collision || domain.enable_excitation_collision ||
                    domain.enable_ionization_collision);
  if (!any) {
    ready_ = false;  // nothing can collide: handle_collisions is a no-op
    return;
  }
  epic_mcc_desc d{};
  d.electron_sp = 0;  // handle_collisions(species_list[0], species_list[1]) (main.cpp:423)
  d.ion_sp = 1;
  d.gas_density = domain.GAS_DENSITY;
  d.gas_temp_K = 0.0;  // neutrals at rest (collision.cpp:63-64)
  d.gas_mass = 0;
  d.e_exc_th = 11.5;
  d.e_ion_th = 15.8;
  d.ntab = NTAB;
  d.e_min = E_MIN;
  d.e_max = E_MAX;
  d.sigma = table_.data();
  d.enable[EPIC_MCC_E_ELASTIC] = domain.enable_elastic_collision;
  d.enable[EPIC_MCC_E_EXCITATION] = domain.enable_excitation_collision;
  d.enable[EPIC_MCC_E_IONIZATION] = domain.enable_ionization_collision;
  d.ref_selection = 1;
  ready_ = true;
  setup_pending_ = true;  // the target mass is known at the first call
  pending_ = d;
}

void CollisionHandler::handle_collisions(Species& electron, Species& target_gas) {
  if (!ready_) return;
  if (setup_pending_) {
    // target mass M = species2.mass (collision.cpp:63-64)
    pending_.electron_sp = electron.index;
    pending_.ion_sp = target_gas.index;
    pending_.gas_mass = target_gas.mass;
    pending_.sigma = table_.data();
    domain.check(epic_mcc_setup(domain.ctx(), &pending_), "epic_mcc_setup");
    setup_pending_ = false;
  }
  domain.check(epic_collide(domain.ctx(), step++), "epic_collide");
}

// ---- Emitter -----------------------------------------------------------------------
Emitter::Emitter(const EmitterParams& p, Domain& domain) : params(p), domain(domain) {
  epic_emitter_desc e{};
  e.x0 = p.x0; e.y0 = p.y0; e.x1 = p.x1; e.y1 = p.y1; e.temp = p.temp;
  e.numparticle = p.numparticle; e.vdx = p.vdx; e.vdy = p.vdy;
  e.species_idx1 = p.species_idx1; e.species_idx2 = p.species_idx2;
  first_ = domain.n_emitters_++ == 0;
  if (domain.ctx_) domain.check(epic_add_emitter(domain.ctx_, &e), "epic_add_emitter");
  else domain.pending_emitters_.push_back(e);
}

// The device injects every emitter in one call (epic_inject); the first
// emitter object issues it once per step, the others have nothing to add.
void Emitter::inject(std::vector<Species>& species_list) {
  (void)species_list;
  epic_ctx* c = domain.ctx();
  if (!first_) return;
  domain.check(epic_inject(c, step++), "epic_inject");
}

// ---- Output (output.cpp:4-315) ---------------------------------------------------
Output::Output(const std::filesystem::path& outputfolder, Domain& domain) : outputfolder(outputfolder), domain(domain) {
  // output.cpp:7-8: the folder is cleared and recreated
  std::filesystem::remove_all(outputfolder);
  std::filesystem::create_directories(outputfolder);
  file = std::make_unique<H5Lite>((outputfolder / "result.h5").string());
  if (!file->ok()) throw std::runtime_error("Error opening HDF5 file");
  field_data_group = file->group(H5Lite::ROOT, "fielddata");
  time_group = file->group(H5Lite::ROOT, "time_var");
  metadata_group = file->group(H5Lite::ROOT, "metadata");
  metadata_species = file->group(H5Lite::ROOT, "metadata_species");
  sp_no = domain.species_no;
  t_step = int(domain.NUM_TS / domain.write_interval) + 1;
  store_ke.assign((size_t)t_step * (3 * sp_no + 2), 0.0);
}

Output::~Output() {
  try {
    if (file) close();
  } catch (const std::exception& e) {
    std::cerr << "Output: " << e.what() << std::endl;
  }
}

void Output::close() {
  if (!file) return;
  const bool good = file->close();
  file.reset();
  if (!good) throw std::runtime_error("Error writing " + (outputfolder / "result.h5").string());
}

static void need(int id, const std::string& what) {
  if (id < 0) throw std::runtime_error("result.h5: cannot create " + what + " (name taken?)");
}

void Output::write_particle_data(int ts, Species& species) {
  const int g = file->group(H5Lite::ROOT, "particle_" + species.name);
  need(g, "group particle_" + species.name);
  std::vector<double> pos, vel;
  species.particles(pos, vel);
  const uint64_t n = pos.size() / 2;
  need(file->dataset_f64(g, "pos" + std::to_string(ts), {n, 2}, pos.data()), "pos" + std::to_string(ts));
  need(file->dataset_f64(g, "vel" + std::to_string(ts), {n, 2}, vel.data()), "vel" + std::to_string(ts));
}

int Output::field_group(const std::string& name) {
  const int g = file->group(field_data_group, name);
  need(g, "group fielddata/" + name);
  return g;
}

// Node (i, j) of an nx x ny field (device layout i*ny + j) to element
// j*nx + i of the dataset, as output.cpp:131-139 and 186-193 fill it.
static std::vector<double> transposed(const std::vector<double>& f, int nx, int ny) {
  std::vector<double> t((size_t)nx * ny);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) t[(size_t)j * nx + i] = f[(size_t)i * ny + j];
  return t;
}

void Output::write_field_data(int ts) {
  const int gp = field_group("pot"), gx = field_group("efieldx"), gy = field_group("efieldy");
  const int nx = domain.nx, ny = domain.ny;
  const std::vector<uint64_t> dims{(uint64_t)nx, (uint64_t)ny};
  const std::string name = std::to_string(ts);
  need(file->dataset_f64(gp, name, dims, transposed(domain.phi(), nx, ny).data()), "pot/" + name);
  need(file->dataset_f64(gx, name, dims, transposed(domain.efx(), nx, ny).data()), "efieldx/" + name);
  need(file->dataset_f64(gy, name, dims, transposed(domain.efy(), nx, ny).data()), "efieldy/" + name);
}

void Output::write_den_data(int ts, Species& species) {
  const int g = field_group("den_" + species.name);
  const int nx = domain.nx, ny = domain.ny;
  const std::string name = std::to_string(ts);
  need(file->dataset_f64(g, name, {(uint64_t)nx, (uint64_t)ny}, transposed(species.den(), nx, ny).data()),
       "den_" + species.name + "/" + name);
}

void Output::write_metadata(int NC, int NUM_TS, int write_int, int write_int_phase, double DT, double density,
                            int save_fig, int normscheme, int subcycleint, double LDe, double LDi, double wpe,
                            double wpi, int spno, double GAS_DENSITY) {
  const int g = metadata_group;
  bool ok = file->attr_i32(g, "NC", NC);
  ok = ok && file->attr_i32(g, "NUM_TS", NUM_TS);
  ok = ok && file->attr_i32(g, "write_int", write_int);
  ok = ok && file->attr_i32(g, "write_int_phase", write_int_phase);
  ok = ok && file->attr_f64(g, "DT_coeff", DT);
  ok = ok && file->attr_f64(g, "density", density);
  ok = ok && file->attr_i32(g, "save_fig", save_fig);
  ok = ok && file->attr_i32(g, "norm_scheme", normscheme);
  ok = ok && file->attr_i32(g, "sub_cycle_interval", subcycleint);
  ok = ok && file->attr_f64(g, "LDe", LDe);
  ok = ok && file->attr_f64(g, "LDi", LDi);
  ok = ok && file->attr_f64(g, "wpe", wpe);
  ok = ok && file->attr_f64(g, "wpi", wpi);
  ok = ok && file->attr_i32(g, "spno", spno);
  ok = ok && file->attr_f64(g, "GAS_DENSITY", GAS_DENSITY);
  if (!ok) throw std::runtime_error("result.h5: metadata written twice");
}

void Output::write_species_metadata(std::vector<Species>& species_list) {
  for (Species& sp : species_list) {
    const int g = file->group(metadata_species, sp.name);
    need(g, "group metadata_species/" + sp.name);
    // output.cpp:239: the name is a PredType::C_S1 attribute, a one-byte
    // string, so it holds the name's first character
    bool ok = file->attr_str(g, "name", sp.name, 1);
    ok = ok && file->attr_f64(g, "mass", sp.mass);
    ok = ok && file->attr_f64(g, "charge", sp.charge);
    ok = ok && file->attr_f64(g, "spwt", sp.spwt);
    ok = ok && file->attr_f64(g, "temperature", sp.temp);
    ok = ok && file->attr_f64(g, "density", sp.fract_den);
    ok = ok && file->attr_i32(g, "num_particles", sp.numparticle);
    ok = ok && file->attr_f64(g, "streaming_velocity", sp.vsx);
    if (!ok) throw std::runtime_error("result.h5: species metadata written twice");
  }
}

void Output::write_ke() {
  const uint64_t cols = (uint64_t)(3 * sp_no + 2);
  need(file->dataset_f64(time_group, "kinetic_energy", {(uint64_t)t_step, cols}, store_ke.data()),
       "time_var/kinetic_energy");
}

void Output::storeKE_to_matrix(int ts, std::vector<Species>& species_list) {
  // output.cpp:294: the normalising species
  Species& norm_species = (domain.SolverType == "hybrid" || domain.SolverType == "gshybrid")
                              ? species_list[0]
                              : ((domain.normscheme == 2 || domain.normscheme == 4) ? species_list[1] : species_list[0]);
  const int k = ts / domain.write_interval;
  if (k < 0 || k >= t_step) throw std::runtime_error("storeKE_to_matrix: step outside the history");
  const size_t cols = (size_t)(3 * sp_no + 2);
  double* row = store_ke.data() + (size_t)k * cols;
  row[0] = ts * domain.DT;
  size_t j = 1;
  for (Species& sp : species_list) {
    const std::array<double, 3> ke = sp.Compute_KE(norm_species);
    row[j] = ke[0];
    row[j + 1] = ke[1];
    row[j + 2] = ke[2];
    j += 3;
  }
  row[j] = domain.Compute_PE(norm_species);
}

}  // namespace epic2d
