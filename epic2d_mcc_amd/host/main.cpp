// epic2d_run: the reference's driver (main.cpp:24-472) on the MI355X path.
// Reads the same input.ini, derives the same normalisation (main.cpp:151-261),
// builds Domain / Species / Emitters, loads particles with the reference's
// seeded loader, and runs NUM_TS+1 steps of the reference's loop
// (main.cpp:395-455).  Diagnostics go to stdout every write_diagnostics steps
// (output.cpp:318-392 without plots); <output>/result.h5 gets the
// reference's metadata, particle, field and density dumps and the energy
// history (main.cpp:268-277, 430-457; output.cpp:4-315), and the energy
// history also goes to <output>/kinetic_energy.csv.  Live plotting is not
// part of this build.
//
// An optional [device] section selects device options:
//   precision = fp64 | fp32, sort_interval = k, tol_relative = 0|1,
//   precond = jacobi | spectral, poisson_op = ref | dirichlet, seed = n
#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "epic_host.hpp"

using namespace epic2d;

static std::string opt(const INIParser::Sections& s, const std::string& sec, const std::string& key,
                       const std::string& def) {
  auto it = s.find(sec);
  if (it == s.end()) return def;
  auto jt = it->second.find(key);
  return jt == it->second.end() ? def : jt->second;
}

int main(int argc, char* argv[]) {
  if (argc < 2) {
    std::cout << "ERROR, at least one argument expected (the input file)." << std::endl;
    return EXIT_FAILURE;
  }
  try {
    auto ini = INIParser::parse(argv[1]);
    auto& species_section = ini["species"];
    const std::string outputfolder = INIParser::getString(ini["file"], "output");
    const int Nx = INIParser::getInt(ini["domain"], "Nx"), Ny = INIParser::getInt(ini["domain"], "Ny");
    const int nx = Nx + 1, ny = Ny + 1;
    const double x0 = INIParser::getDouble(ini["domain"], "x0");
    const std::string elastic_flag = INIParser::getString(ini["collision"], "elastic");
    const std::string excitation_flag = INIParser::getString(ini["collision"], "excitation");
    const std::string ionization_flag = INIParser::getString(ini["collision"], "ionization");
    const double GAS_DENSITY = INIParser::getDouble(ini["collision"], "GAS_DENSITY");
    const int NUM_TS = INIParser::getInt(ini["time"], "NUM_TS");
    const double DT_coeff = INIParser::getDouble(ini["time"], "DT_coeff");
    const std::string shapefunction = INIParser::getString(ini["simulation"], "shapefunction");
    const double den = INIParser::getDouble(ini["simulation"], "density");
    const std::string bc = INIParser::getString(ini["simulation"], "bc");
    const int ionfixed = INIParser::getInt(ini["simulation"], "ionfixed");
    const int precision = INIParser::getInt(ini["diagnostics"], "precision");
    const int write_interval = INIParser::getInt(ini["diagnostics"], "write_interval");
    const int write_interval_phase = INIParser::getInt(ini["diagnostics"], "write_interval_phase");
    const int write_diagnostics = INIParser::getInt(ini["diagnostics"], "write_diagnostics");
    const int write_flag = INIParser::getInt(ini["diagnostics"], "write_flag");
    const int save_fig = INIParser::getInt(ini["diagnostics"], "save_fig");
    const int sub_cycle_interval = INIParser::getInt(ini["diagnostics"], "sub_cycle_interval");
    const int norm_scheme = INIParser::getInt(ini["normalization"], "norm_scheme");
    const double L_scale = INIParser::getDouble(ini["normalization"], "lenght_scale");
    const std::string T_scale = INIParser::getString(ini["normalization"], "time_scale");
    const std::string SolverType = INIParser::getString(ini["solver"], "solvertype");
    const double tolerance = INIParser::getDouble(ini["solver"], "tolerance");
    const int max_iteration = INIParser::getInt(ini["solver"], "max_iteration");

    std::vector<std::string> names, pos_init;
    std::vector<double> mass, temps, frac, vsx, vsy;
    std::vector<int> nParticles, signs;
    for (auto& e : species_section) {
      auto t = INIParser::split(e.second, ',');
      if (t.size() != 9) continue;
      names.push_back(t[0]);
      mass.push_back(std::stod(t[1]));
      nParticles.push_back(std::stoi(t[2]));
      temps.push_back(std::stod(t[3]));
      signs.push_back(std::stoi(t[4]));
      frac.push_back(std::stod(t[5]));
      vsx.push_back(std::stod(t[6]));
      vsy.push_back(std::stod(t[7]));
      pos_init.push_back(t[8]);
    }
    const int species_no = (int)names.size();
    if (species_no < 2) throw std::runtime_error("at least two species are needed (main.cpp:159-160)");
    // main.cpp:151-172
    double k = 0;
    for (int i = 0; i < species_no; i++) k += (-signs[i]) * frac[i];
    std::vector<double> normden(species_no);
    normden[1] = den;
    normden[0] = den / k;
    for (int i = 2; i < species_no; i++) normden[i] = frac[i] * normden[0];
    using namespace Const;
    const double LDe = std::sqrt((EPS_0 * K_b * temps[0] * EV_to_K) / (normden[0] * QE * QE));
    const double wpe = std::sqrt((normden[0] * QE * QE) / (mass[0] * EPS_0));
    const double wpi = std::sqrt((normden[1] * QE * QE) / (mass[1] * EPS_0));
    const double LDi = std::sqrt((EPS_0 * K_b * temps[1] * EV_to_K) / (normden[1] * QE * QE));
    const double CS = std::sqrt(temps[0] * K_b * EV_to_K / mass[1]);
    double time_scale = 0;
    if (T_scale == "omegape") time_scale = wpe;
    else if (T_scale == "omegapi") time_scale = wpi;
    else time_scale = std::stod(T_scale);
    double dx = 0, dy = 0, DT = 0, stepSize;  // main.cpp:196-224
    if (norm_scheme == 2 || norm_scheme == 4) {
      stepSize = LDi; dx = stepSize / LDi; dy = stepSize / LDi;
      DT = DT_coeff * (1.0 / wpi); DT = wpi * DT;
    }
    if (norm_scheme == 1 || norm_scheme == 3) {
      stepSize = LDe; dx = stepSize / LDe; dy = stepSize / LDe;
      DT = DT_coeff * (1.0 / wpe); DT = wpe * DT;
    }
    if (norm_scheme == 5) {
      stepSize = L_scale; dx = stepSize / L_scale; dy = stepSize / L_scale;
      DT = DT_coeff * (1.0 / time_scale); DT = time_scale * DT;
    }
    Domain domain(x0, x0, dx, dy, nx, ny);  // main.cpp:226 (y0 := x0)
    domain.bc = bc;
    domain.SolverType = SolverType;
    domain.tolerance = tolerance;
    domain.species_no = species_no;
    domain.density = den;
    domain.shapefunction = shapefunction;
    domain.IAW_vel = CS;
    domain.max_iteration = max_iteration;
    domain.normscheme = norm_scheme;
    domain.sub_cycle_interval = sub_cycle_interval;
    domain.lenght_scale = L_scale;
    domain.time_scale = time_scale;
    domain.set_normparam(LDe, wpe, LDi, wpi);
    domain.set_time(DT, NUM_TS, write_interval, write_interval_phase);
    domain.set_normscheme();
    domain.ionfixed = ionfixed;
    domain.enable_elastic_collision = string_to_bool(elastic_flag);
    domain.enable_excitation_collision = string_to_bool(excitation_flag);
    domain.enable_ionization_collision = string_to_bool(ionization_flag);
    domain.GAS_DENSITY = GAS_DENSITY;
    // device options
    domain.precision = opt(ini, "device", "precision", "fp64") == "fp32" ? EPIC_FP32 : EPIC_FP64;
    domain.sort_interval = std::stoi(opt(ini, "device", "sort_interval", "0"));
    domain.tol_relative = std::stoi(opt(ini, "device", "tol_relative", "0"));
    domain.precond = opt(ini, "device", "precond", "jacobi") == "spectral" ? EPIC_PRECOND_SPECTRAL
                                                                            : EPIC_PRECOND_JACOBI;
    const std::string op = opt(ini, "device", "poisson_op", "");
    if (op == "ref") domain.poisson_op = EPIC_OP_REF_GHOST;
    if (op == "dirichlet") domain.poisson_op = EPIC_OP_DIRICHLET_WALL;
    domain.seed = std::stoull(opt(ini, "device", "seed", "1"));
    domain.capacity_factor = std::stod(opt(ini, "device", "capacity_factor", "1.5"));

    std::vector<Species> species_list;
    species_list.reserve(species_no);
    for (int i = 0; i < species_no; i++) {
      const double spwt = normden[i] * domain.Lx * domain.Ly * domain.L * domain.L / nParticles[i];  // main.cpp:256
      species_list.emplace_back(names[i], mass[i], signs[i] * QE, spwt, temps[i], nParticles[i], vsx[i], vsy[i],
                                frac[i], pos_init[i], domain);
    }
    domain.bind_species(species_list);

    // main.cpp:287-323: conductors, voltages normalised by kT/e at 1 eV
    std::vector<Grid*> grids;
    const auto& gs = ini.at("Grids");
    const int grid_count = INIParser::getInt(gs, "grid_count");
    for (int i = 0; i < grid_count; i++) {
      const std::string p = "grid_" + std::to_string(i) + ".";
      const std::string type = INIParser::getString(gs, p + "type");
      GridParams gp{};
      RectangularGridParams rp{};
      if (type == "circular") {
        gp.grid_centerx = INIParser::getInt(gs, p + "x_center");
        gp.grid_centery = INIParser::getInt(gs, p + "y_center");
        gp.num_electrodes = INIParser::getInt(gs, p + "electrode_number");
        gp.grid_rad = INIParser::getDouble(gs, p + "grid_radius");
        gp.electrode_rad = INIParser::getDouble(gs, p + "electrod_radius");
        gp.voltage = (Const::eV / (Const::K_b * Const::EV_to_K)) * INIParser::getDouble(gs, p + "electrode_voltage");
        grids.push_back(Create_Grid(GridType::Circular, domain, gp, rp));
      } else if (type == "reactconduct") {
        rp.min_x = INIParser::getInt(gs, p + "min_x");
        rp.min_y = INIParser::getInt(gs, p + "min_y");
        rp.max_x = INIParser::getInt(gs, p + "max_x");
        rp.max_y = INIParser::getInt(gs, p + "max_y");
        rp.voltage = (Const::eV / (Const::K_b * Const::EV_to_K)) * INIParser::getDouble(gs, p + "grid_voltage");
        grids.push_back(Create_Grid(GridType::ReactConductor, domain, gp, rp));
      }
    }
    std::vector<Emitter> emitters;
    const auto& es = ini.at("Emitters");
    const int emitter_count = INIParser::getInt(es, "count");
    emitters.reserve(emitter_count);
    for (int i = 0; i < emitter_count; i++) {
      const std::string p = "emitter_" + std::to_string(i) + ".";
      EmitterParams ep;
      ep.x0 = INIParser::getDouble(es, p + "x0");
      ep.y0 = INIParser::getDouble(es, p + "y0");
      ep.x1 = INIParser::getDouble(es, p + "x1");
      ep.y1 = INIParser::getDouble(es, p + "y1");
      ep.temp = INIParser::getDouble(es, p + "temp");
      ep.numparticle = INIParser::getInt(es, p + "numparticle");
      ep.vdx = INIParser::getDouble(es, p + "vdx");
      ep.vdy = INIParser::getDouble(es, p + "vdy");
      ep.species_idx1 = INIParser::getInt(es, p + "species_idx1");
      ep.species_idx2 = INIParser::getInt(es, p + "species_idx2");
      emitters.emplace_back(ep, domain);
    }

    // main.cpp:268-277 (the output folder is cleared and recreated)
    Output output(outputfolder, domain);
    output.precision = precision;
    output.write_metadata(nx, NUM_TS, write_interval, write_interval_phase, DT_coeff, den, save_fig,
                          domain.normscheme, domain.sub_cycle_interval, LDe, LDi, wpe, wpi, species_no, 0);
    output.write_species_metadata(species_list);

    const auto start_time = std::chrono::high_resolution_clock::now();
    for (Species& sp : species_list) Init init(sp, domain, grids);
    FieldSolve fieldsolver(domain, grids);
    for (Species& sp : species_list) sp.ScatterSpecies_serial();
    domain.ComputeRho(species_list);
    fieldsolver.PotentialSolver();
    fieldsolver.CalculateEfield();
    for (Species& sp : species_list) sp.Rewind_species();
    CollisionHandler testcoll(domain);
    testcoll.set_electron_cross_sections();
    testcoll.calc_total_cross_sections();

    // output.cpp:323-324: the normalising species
    Species& norm_species = (SolverType == "hybrid" || SolverType == "gshybrid")
                                ? species_list[0]
                                : ((norm_scheme == 2 || norm_scheme == 4) ? species_list[1] : species_list[0]);
    std::ofstream ke_out(std::filesystem::path(outputfolder) / "kinetic_energy.csv");
    ke_out << std::setprecision(17);
    for (int ts = 0; ts < NUM_TS + 1; ts++) {
      for (auto& emit : emitters) emit.inject(species_list);
      for (Species& sp : species_list) sp.ScatterSpecies_serial();
      domain.ComputeRho(species_list);
      fieldsolver.PotentialSolver();
      fieldsolver.CalculateEfield();
      for (Species& sp : species_list) {
        if (sp.name == "ion" && domain.ionfixed == 1) continue;
        sp.Push_species_serial(1, grids);
      }
      testcoll.handle_collisions(species_list[0], species_list[1]);
      if (ts % write_diagnostics == 0) {
        const auto phi = domain.phi();
        double max_phi = phi[0];
        for (double v : phi) max_phi = std::max(max_phi, v);
        std::cout << "TS: " << ts << "\tnorm: " << domain.error << " delta_phi: " << std::fixed
                  << std::setprecision(2) << (max_phi - phi[0]);
        double ke_tot = 0;
        std::array<double, 3> p{0, 0, 0};
        for (Species& sp : species_list) {
          auto ke = sp.Compute_KE(norm_species);
          ke_tot += ke[0] + ke[1] + ke[2];
          auto m = sp.Compute_Momentum(norm_species);
          for (int c = 0; c < 3; c++) p[c] += m[c];
          if (bc == "open") std::cout << " n_" << sp.name << ":" << sp.count();
          if (bc == "pbc")
            std::cout << " KE_" << sp.name << ": " << std::setprecision(precision) << ke[0] + ke[1] + ke[2];
        }
        const double pe = domain.Compute_PE(norm_species);
        std::cout << " Potential_energy: " << std::setprecision(precision) << pe << " Kinetic_energy: " << ke_tot
                  << " Total_energy: " << ke_tot + pe;
        if (bc == "pbc") std::cout << " p_x: " << p[0] << " p_y: " << p[1] << " p_z: " << p[2];
        std::cout << std::defaultfloat << std::endl;
      }
      // main.cpp:432-453
      if (ts % write_interval_phase == 0 && ts % write_interval == 0)
        for (Species& sp : species_list) output.write_particle_data(ts, sp);
      if (ts % write_interval == 0) {
        output.storeKE_to_matrix(ts, species_list);
        const size_t cols = 3 * species_list.size() + 2;
        const double* row = output.store_ke.data() + (size_t)(ts / write_interval) * cols;
        for (size_t c = 0; c < cols; c++) ke_out << (c ? "," : "") << row[c];
        ke_out << "\n";
        if (write_flag == 1 || write_flag == 2) {
          output.write_field_data(ts);
          output.write_den_data(ts, species_list[0]);
          output.write_den_data(ts, species_list[1]);
        }
      }
    }
    output.write_ke();  // main.cpp:457
    output.close();
    const auto end_time = std::chrono::high_resolution_clock::now();
    const double el = std::chrono::duration<double>(end_time - start_time).count();
    std::cout << "Elapsed time: " << el << " seconds" << std::endl;
    for (Grid* g : grids) delete g;  // main.cpp:460-463
  } catch (const std::exception& e) {
    std::cerr << "epic2d_run: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
  return 0;
}
