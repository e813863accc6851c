// Host C++ API with the reference's class and method names (Domain, Species,
// Init, FieldSolve, CollisionHandler, Emitter, INIParser), so that the
// reference's main loop (main.cpp:366-455) keeps its shape.  Every hot
// method forwards to the C-ABI in include/epic2d.h, which runs gfx950 HIP
// kernels; particles and fields live on the device.
#pragma once

#include <array>
#include <filesystem>
#include <map>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "epic2d.h"
#include "h5lite.hpp"

namespace epic2d {

namespace Const {  // include/domain.h:204-215
const double EPS_0 = 8.85418782e-12;
const double QE = 1.602176565e-19;
const double AMU = 1.660538921e-27;
const double ME = 9.10938215e-31;
const double K_b = 1.380648e-23;
const double PI = 3.141592653;
const double EV_to_K = QE / K_b;
const double eV = 1.602176565e-19;
}  // namespace Const

// The reference's global random stream (domain.h:218-231): mt19937 seeded 0
// with uniform_real_distribution<double>(0, 1).  Init draws from it.
class Rnd {
 public:
  Rnd() : mt_gen{0}, rnd_dist{0, 1.0} {}
  double operator()() { return rnd_dist(mt_gen); }
  void seed(unsigned s) { mt_gen.seed(s); }

 private:
  std::mt19937 mt_gen;
  std::uniform_real_distribution<double> rnd_dist;
};
extern Rnd rnd;

// Failure of a C-ABI call: status code and the library's message.
class EpicError : public std::runtime_error {
 public:
  EpicError(int code, const std::string& what) : std::runtime_error(what), code(code) {}
  int code;
};

// INIParser (include/iniparser.h:20-39, src/iniparser.cpp:4-113).
class INIParser {
 public:
  using Sections = std::map<std::string, std::map<std::string, std::string>>;
  static Sections parse(const std::string& filename);
  static Sections parse_text(const std::string& text);
  static int getInt(const std::map<std::string, std::string>& section, const std::string& key);
  static double getDouble(const std::map<std::string, std::string>& section, const std::string& key);
  static std::string getString(const std::map<std::string, std::string>& section, const std::string& key);
  static std::vector<std::string> split(const std::string& str, char delimiter);
  static std::pair<std::string, int> loadtypeextract(const std::string& position_init);

 private:
  static void trim(std::string& str);
};

bool string_to_bool(const std::string& s);  // include/extrafun.h:6-17

class Species;
class Grid;

class Domain {
 public:
  double x0, y0, Lx, Ly, dx, dy;
  int nx, ny;
  int NUM_TS = 0, write_interval = 1, write_interval_phase = 1, sub_cycle_interval = 1;
  double LDe = 0, wpe = 0, LDi = 0, wpi = 0, lenght_scale = 0, time_scale = 0, W = 0, L = 0, IAW_vel = 0;
  double vel_norm = 0, density = 0, DT = 0;
  int species_no = 0;
  std::string SolverType = "cg";
  double tolerance = 1e-8;
  int max_iteration = 10000;
  double error = 0;
  int last_iterations = 0;
  std::string bc = "pbc";
  std::string shapefunction = "CIC";
  int normscheme = 1;
  int vel_normscheme = 0;
  int ionfixed = 0;
  bool enable_elastic_collision = false, enable_excitation_collision = false, enable_ionization_collision = false;
  double GAS_DENSITY = 0;
  // device options (no reference counterpart)
  int precision = EPIC_FP64;
  int tol_relative = 0;
  int sort_interval = 0;
  int poisson_op = -1;  // -1: the reference's operator for the solver type
  int precond = EPIC_PRECOND_JACOBI;
  unsigned long long seed = 1;
  double capacity_factor = 1.0;

  Domain(double x0, double y0, double dx, double dy, int nx, int ny);
  ~Domain();
  Domain(const Domain&) = delete;
  Domain& operator=(const Domain&) = delete;

  void set_normparam(double LDe, double wpe, double LDi, double wpi);
  void set_normscheme();
  void set_time(double DT, int NUM_TS, int write_interval, int write_interval_phase);
  void ComputeRho(std::vector<Species>& species);
  double Compute_PE(Species& species);
  double XtoL(double pos) const { return (pos - x0) / dx; }
  double YtoL(double pos) const { return (pos - y0) / dy; }

  // field access (nx*ny doubles, (i, j) -> i*ny + j)
  std::vector<double> rho() { return field(0); }
  std::vector<double> phi() { return field(1); }
  std::vector<double> efx() { return field(2); }
  std::vector<double> efy() { return field(3); }
  std::vector<double> field(int which);

  // device context (created when the first species needs the device)
  epic_ctx* ctx();
  void check(int rc, const char* what);
  int register_species(Species* s);
  void bind_species(std::vector<Species>& list);
  // the conductors the device pushes and solves with (epic_set_conductors);
  // a repeated call with the same list does nothing
  void set_grids(const std::vector<Grid*>& grids);

 private:
  epic_ctx* ctx_ = nullptr;
  std::vector<Grid*> grids_;
  bool grids_set_ = false;
  std::vector<Species*> reg_;
  std::vector<epic_emitter_desc> pending_emitters_;
  int n_emitters_ = 0;
  friend class Emitter;
};

class Species {
 public:
  std::string name;
  double mass, charge, spwt, temp;
  int numparticle;
  int charge_sig = 0;
  std::string initialization;
  double vsx, vsy, fract_den;
  int index;

  Species(std::string name, double mass, double charge, double spwt, double temp, int numparticle, double vsx,
          double vsy, double fract_den, std::string initialization, Domain& domain);

  void Push_species_serial(int sub_cycle, const std::vector<Grid*>& grids);  // species.cpp:29-170
  void ScatterSpecies_serial();                                              // species.cpp:172-226
  void Rewind_species();                                                     // species.cpp:229-246
  std::array<double, 3> Compute_KE(Species& species);                        // species.cpp:248-271
  std::array<double, 3> Compute_Momentum(Species& species);                  // species.cpp:273-297
  int64_t count();
  std::vector<double> den();
  // The live particles in storage order as interleaved (x, y) and (vx, vy)
  // pairs: Particle::pos[0..1], vel[0..1] of part_list (species.h:13-20).
  void particles(std::vector<double>& pos, std::vector<double>& vel);
  Domain& domain() { return *domain_; }

 private:
  Domain* domain_;
};

// ---- internal conductors (grid.h:13-88, electrode.h:9-20) ----------------------
// The host keeps each grid's geometry, tests it for the loader (IsGrid) and
// hands it to the device (desc()), which absorbs particles in the push and
// writes the voltages before each SOR solve.
enum class GridType { Custom, Circular, Line, ReactConductor };

struct GridParams {  // grid.h:69-77
  int grid_centerx;
  int grid_centery;
  int num_electrodes;
  double grid_rad;
  double electrode_rad;
  double voltage;
};

struct RectangularGridParams {  // grid.h:79-86
  int min_x;
  int max_x;
  int min_y;
  int max_y;
  double voltage;
};

class Grid {
 public:
  explicit Grid(Domain& domain) : domain(domain) {}
  virtual ~Grid() = default;
  virtual bool IsGrid(double x, double y) = 0;
  virtual void initialize() = 0;
  virtual epic_conductor_desc desc() const = 0;

 protected:
  Domain& domain;
};

// A ring of electrode disks; they exist once initialize() has run (an SOR
// solve), as in grid.cpp:24-50.
class CircularGrid : public Grid {
 public:
  CircularGrid(Domain& domain, int center_x, int center_y, int num_electrodes, double voltage,
               double electrode_rad, double grid_rad);
  bool IsGrid(double x, double y) override;
  void initialize() override;
  epic_conductor_desc desc() const override;

 private:
  int center_x, center_y, num_electrodes;
  double voltage, electrode_rad, grid_rad;
  std::vector<double> ex_, ey_;  // electrode centres after initialize()
};

class ReactConductorGrid : public Grid {  // grid.cpp:53-85
 public:
  ReactConductorGrid(int min_x, int max_x, int min_y, int max_y, double voltage, Domain& domain);
  bool IsGrid(double x, double y) override;
  void initialize() override {}  // the device writes the boundary voltages
  epic_conductor_desc desc() const override;

 private:
  int min_x, max_x, min_y, max_y;
  double voltage;
};

// grid.cpp:88-103; GridType::Custom (electrodes added one by one) is not
// supported and throws.
Grid* Create_Grid(GridType type, Domain& domain, const GridParams& params, const RectangularGridParams& rparam);

// Init::Init (init.cpp:3-102): the reference's loader, drawing from the
// global mt19937 stream in the reference's order, then uploading.
class Init {
 public:
  Init(Species& species, Domain& domain, std::vector<Grid*>& grids);
  static double SampleVel(Species& species);
  static double SampleVel(Species& species, double temp);
};

class FieldSolve {
 public:
  FieldSolve(Domain& domain, const std::vector<Grid*>& grids);
  void PotentialSolver();  // field.cpp:266-272
  void CalculateEfield();  // field.cpp:276-331

 private:
  Domain& domain;
  std::vector<Grid*> grids;
};

class CollisionHandler {
 public:
  explicit CollisionHandler(Domain& domain) : domain(domain) {}
  // Tabulate the reference's analytic e-Ar cross sections
  // (cross_section.cpp:4-39) for the enabled processes and set up the
  // null-collision MCC (collision.cpp:21-42).
  void set_electron_cross_sections();
  void calc_total_cross_sections();
  void handle_collisions(Species& electron, Species& target_gas);  // collision.cpp:190-207
  int step = 0;

 private:
  Domain& domain;
  std::vector<double> table_;
  bool ready_ = false;
  bool setup_pending_ = false;
  epic_mcc_desc pending_{};
};

struct EmitterParams {  // emitter.h:8-20
  double x0, y0, x1, y1, temp;
  int numparticle;
  double vdx, vdy;
  int species_idx1, species_idx2;
};

class Emitter {
 public:
  Emitter(const EmitterParams& params, Domain& domain);
  void inject(std::vector<Species>& species_list);  // emitter.cpp:5-36
  int step = 0;

 private:
  EmitterParams params;
  Domain& domain;
  bool first_ = false;
};

// Output (output.h:37-101): result.h5 in the reference's layout, written by
// the self-contained HDF5 writer of h5lite.hpp.
//   /metadata                       scalar attributes (output.cpp:202-225)
//   /metadata_species/<name>        scalar attributes (output.cpp:228-249)
//   /particle_<name>/pos<ts>, vel<ts>   N x 2 doubles (output.cpp:33-78)
//   /fielddata/{pot,efieldx,efieldy}/<ts>, /fielddata/den_<name>/<ts>
//       nx x ny doubles holding the field transposed: element j*nx + i of
//       the row-major data is node (i, j) (output.cpp:81-199)
//   /time_var/kinetic_energy        (NUM_TS/write_interval + 1) x (3 nsp + 2):
//       t, KE (x, y, z) per species, PE (output.cpp:253-315)
// The file is complete when the Output is destroyed (or close() returns).
class Output {
 public:
  Output(const std::filesystem::path& outputfolder, Domain& domain);  // output.cpp:4-31
  ~Output();
  Output(const Output&) = delete;
  Output& operator=(const Output&) = delete;

  void write_particle_data(int ts, Species& species);  // output.cpp:33-78
  void write_field_data(int ts);                       // output.cpp:81-147
  void write_den_data(int ts, Species& species);       // output.cpp:150-199
  void write_metadata(int NC, int NUM_TS, int write_int, int write_int_phase, double DT, double density, int save_fig,
                      int normscheme, int subcycleint, double LDe, double LDi, double wpe, double wpi, int spno,
                      double GAS_DENSITY);             // output.cpp:202-225
  void write_species_metadata(std::vector<Species>& species_list);     // output.cpp:228-249
  void write_ke();                                                     // output.cpp:253-288
  void storeKE_to_matrix(int ts, std::vector<Species>& species_list);  // output.cpp:291-315
  // Writes the object headers and the superblock; throws on an I/O error.
  void close();

  int precision = 6;
  std::vector<double> store_ke;  // t_step x (3 sp_no + 2), row-major
  int sp_no = 0, t_step = 0;

 private:
  int field_group(const std::string& name);
  std::filesystem::path outputfolder;
  Domain& domain;
  std::unique_ptr<H5Lite> file;
  int field_data_group = -1, time_group = -1, metadata_group = -1, metadata_species = -1;
};

// Analytic e-Ar cross sections of the reference (cross_section.cpp:4-39).
double compute_elastic_CS(double energy);
double compute_excitation_CS(double energy);
double compute_ionization_CS(double energy);

}  // namespace epic2d
