// C++ host layer above libkomd_hip.so (see komd_host.hpp).
#include "komd_host.hpp"

#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>

namespace komd_host {

void check(komd_ctx* ctx, komd_status st, const char* what) {
    if (st == KOMD_OK) return;
    std::string msg = std::string(what) + ": ";
    msg += ctx ? komd_last_error(ctx) : "invalid argument";
    throw std::runtime_error(msg);
}

// ---------------------------------------------------------------------------
// deck (IO/FileInput.cpp:11-121): whitespace-separated "key values" pairs, '#' starts a
// comment line; keys not used by the force path are read past.
bool readFile(const std::string& path, Configuration& config, std::vector<Component>& components, std::string* err) {
    auto fail = [&](const std::string& m) {
        if (err) *err = m;
        return false;
    };
    std::ifstream file(path);
    if (!file.is_open()) return fail("Could not open the file: " + path);
    double max_eps = std::numeric_limits<double>::min(), max_sig = std::numeric_limits<double>::min();
    std::string var;
    auto d3in = [&](d3& x) { file >> x[0] >> x[1] >> x[2]; };
    while (file >> var) {
        if (var[0] == '#') {
            file.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
            continue;
        }
        if (var == "temperature") file >> config.temperature;
        else if (var == "delta_t") file >> config.delta_t;
        else if (var == "cutoff") file >> config.cutoff;
        else if (var == "cell_size") file >> config.cell_size;
        else if (var == "limit") file >> config.limit_factor;
        else if (var == "energy_3b") file >> config.energy_3b;
        else if (var == "domain_low") d3in(config.domainLow);
        else if (var == "domain_high") d3in(config.domainHigh);
        else if (var == "timesteps") file >> config.timesteps;
        else if (var == "write_freq") file >> config.write_freq;
        else if (var == "enable_3b") file >> config.enable_3b;
        else if (var == "enable_3b_direct") file >> config.enable_3b_direct;
        else if (var == "pair_steps") file >> config.pair_steps;
        else if (var == "triple_steps") file >> config.triple_steps;
        else if (var == "store_checkpoint") file >> config.storeCheckpoint;
        else if (var == "checkpoint_file") {
            std::string p;
            d3 off;
            file >> p;
            d3in(off);
            config.checkpoint_files.emplace_back(p, off);
        } else if (var == "COMP") {
            int cid;
            Component::Site s;
            file >> cid >> s.eps >> s.sig >> s.mass >> s.r[0] >> s.r[1] >> s.r[2];
            if (cid < 0) return fail("negative component id not allowed");
            if ((int)components.size() <= cid) components.resize((size_t)cid + 1);
            components[(size_t)cid].sites.push_back(s);
            max_eps = std::max(max_eps, s.eps);
            max_sig = std::max(max_sig, s.sig);
        }
        // GPU path keys
        else if (var == "atm_mode") file >> config.atm_mode;
        else if (var == "surrogate_weights") file >> config.surrogate_weights;
        else if (var == "judge_every") file >> config.judge_every;
        else if (var == "judge_max") file >> config.judge_max;
        else if (var == "hybrid_region") file >> config.hybrid_axis >> config.hybrid_center >> config.hybrid_half >> config.hybrid_blend;
        else if (var == "thermostat") file >> config.thermostat;
        else if (var == "skin") file >> config.skin;
        else if (var == "checkpoint_precision") file >> config.checkpoint_precision;
        else if (var == "enable_3b_approx") file >> config.enable_3b_approx;
        else if (var == "pair_table") file >> config.pair_table;
        else if (var == "pair_table_potential") file >> config.pair_table_potential;
        else if (var == "enable_sensor_rdf") file >> config.enable_sensor_rdf;
        else if (var == "enable_sensor_pres") file >> config.enable_sensor_pres;
        else if (var == "sensor_rdf_max") file >> config.sensor_rdf_max;
        else if (var == "sensor_rdf_dr") file >> config.sensor_rdf_dr;
        else if (var == "sensor_dir") file >> config.sensor_dir;
        if (file.fail()) return fail("malformed value for key " + var);
    }
    if (components.empty()) return fail("no components defined");
    if (max_eps == std::numeric_limits<double>::min()) return fail("invalid epsilon specified");
    if (max_sig == std::numeric_limits<double>::min()) return fail("invalid sigma specified");
    if (config.checkpoint_files.empty()) return fail("No phasespace data specified");
    if (config.cell_size < config.cutoff) return fail("cell size must be larger or equal to cutoff radius");
    if (config.atm_mode != "reference" && config.atm_mode != "exact") return fail("atm_mode: reference or exact");
    if (config.pair_steps < 1 || config.triple_steps < 1) return fail("pair_steps and triple_steps must be positive");
    if (!(config.skin >= 0)) return fail("skin must be non-negative");
    if (config.skin > 0 && config.cell_size < config.cutoff + config.skin)
        return fail("cell size must be at least cutoff + skin");
    return true;
}

void Sites::add(uint64_t mol_id, const d3& pos, const d3& vel, double m, double e, double s) {
    r.insert(r.end(), pos.begin(), pos.end());
    v.insert(v.end(), vel.begin(), vel.end());
    f.insert(f.end(), 3, 0.0);
    mass.push_back(m);
    eps.push_back(e);
    sig.push_back(s);
    id.push_back(mol_id);
}

// ---------------------------------------------------------------------------
// checkpoints (IO/CheckpointIO.cpp:13-76): "id  rx ry rz  vx vy vz  mass eps sig" per
// site; consecutive lines with one id form one molecule; epsilon is internal.
namespace CheckpointIO {

void readPs(const std::string& path, const d3& offset, Sites& sites) {
    std::ifstream file(path);
    if (!file.is_open()) throw std::runtime_error("Could not open checkpoint file: " + path);
    std::string line;
    size_t lineno = 0;
    while (std::getline(file, line)) {
        lineno++;
        std::istringstream in(line);
        uint64_t id;
        d3 r, v;
        double m, e, s;
        if (!(in >> id)) continue;  // blank line
        if (!(in >> r[0] >> r[1] >> r[2] >> v[0] >> v[1] >> v[2] >> m >> e >> s))
            throw std::runtime_error(path + ":" + std::to_string(lineno) + ": malformed site line");
        for (int k = 0; k < 3; k++) r[k] += offset[k];
        sites.add(id, r, v, m, e, s);
    }
}

void loadCheckpoint(const Configuration& config, Sites& sites) {
    for (const auto& cf : config.checkpoint_files) readPs(cf.first, cf.second, sites);
}

void writePs(const std::string& path, const Sites& s, int precision) {
    std::ofstream file(path);
    if (!file.is_open()) throw std::runtime_error("Could not write checkpoint file: " + path);
    file << std::setprecision(precision);
    for (size_t i = 0; i < s.size(); i++) {
        if (i) file << "\n";
        file << s.id[i] << "\t" << s.r[3 * i] << " " << s.r[3 * i + 1] << " " << s.r[3 * i + 2] << "\t" << s.v[3 * i]
             << " " << s.v[3 * i + 1] << " " << s.v[3 * i + 2] << "\t" << s.mass[i] << " " << s.eps[i] << " "
             << s.sig[i];
    }
}

void writeCheckpoint(uint64_t simstep, MoleculeContainer& container, int precision, const std::string& path) {
    container.writeSOA2AOS();
    writePs(path.empty() ? "ckpt_" + std::to_string(simstep) + ".ps" : path, container.sites(), precision);
}

}  // namespace CheckpointIO

// ---------------------------------------------------------------------------
MoleculeContainer::~MoleculeContainer() {
    if (m_ctx) komd_destroy(m_ctx);
}

void MoleculeContainer::constructSOAs() {
    const size_t n = m_sites.size();
    check(m_ctx, komd_set_sites(m_ctx, (int64_t)n, m_sites.r.data(), m_sites.v.data(), m_sites.mass.data(),
                                m_sites.eps.data(), m_sites.sig.data(), m_sites.id.data()),
          "constructSOAs");
    m_num_molecules = 0;
    for (size_t i = 0; i < n; i++)
        if (i == 0 || m_sites.id[i] != m_sites.id[i - 1]) m_num_molecules++;
}

void MoleculeContainer::writeSOA2AOS() {
    check(m_ctx, komd_get_sites(m_ctx, m_sites.r.data(), m_sites.v.data(), m_sites.f.data()), "writeSOA2AOS");
}

LinkedCells::LinkedCells(const Configuration& config) {
    for (int k = 0; k < 3; k++) {
        m_box.low[k] = config.domainLow[k];
        m_box.high[k] = config.domainHigh[k];
    }
    m_box.cell_size = config.cell_size;
    m_params.cutoff = config.cutoff;
    m_params.delta_t = config.delta_t;
    m_params.limit_factor = config.limit_factor;
    m_params.nu = config.energy_3b * 1e9 * kConvJEi;  // ATM.cpp:13
    m_params.temp_factor = kTemperatureFactor;
    m_params.bin_mode = KOMD_BIN_COM;
}

void LinkedCells::init() {
    komd_ctx* c = nullptr;
    const komd_status st = komd_create(&m_box, &m_params, &c);
    m_ctx = c;
    check(m_ctx, st, "LinkedCells::init");
    constructSOAs();
}

void LinkedCells::updateContainer() { check(m_ctx, komd_update_container(m_ctx), "updateContainer"); }

std::array<int64_t, 3> LinkedCells::gridDims() const {
    std::array<int64_t, 3> d{};
    check(m_ctx, komd_grid_dims(m_ctx, d.data()), "gridDims");
    return d;
}

void LJ12_6::handlePairList() { check(p_container.ctx(), komd_forces_pair(p_container.ctx(), KOMD_PAIR_LJ12_6), "LJ12_6"); }
void ATM2B::handlePairList() { check(p_container.ctx(), komd_forces_pair(p_container.ctx(), KOMD_PAIR_ATM2B), "ATM2B"); }

double FunctionPL::evaluateAt(double x) const {
    const size_t n = m_x.size();
    if (n == 0) throw std::runtime_error("FunctionPL: empty function");
    if (x > m_x[n - 1]) return m_high;
    if (x < m_x[0]) return m_low;
    size_t i = 0;  // findUpperNode: the first node >= x
    while (x > m_x[i]) {
        i++;
        if (i == n) break;
    }
    if (i == 0) return m_y[0];
    const double ya = m_y[i - 1], yb = m_y[i], xa = m_x[i - 1], xb = m_x[i];
    return ya + (yb - ya) / (xb - xa) * (x - xa);
}

void FunctionPL::derivative(FunctionPL& target) const {
    const size_t n = m_y.size();
    if (n < 2) throw std::runtime_error("FunctionPL::derivative: fewer than two nodes");
    target.m_x = m_x;
    target.m_y.assign(n, 0.0);
    target.m_y[0] = (m_y[1] - m_y[0]) / (m_x[1] - m_x[0]);
    target.m_y[n - 1] = (m_y[n - 1] - m_y[n - 2]) / (m_x[n - 1] - m_x[n - 2]);
    for (size_t k = 1; k + 1 < n; k++) target.m_y[k] = (m_y[k + 1] - m_y[k - 1]) / (m_x[k + 1] - m_x[k - 1]);
}

void FunctionPL::read(const std::string& path) {
    std::ifstream file{path};
    if (!file) throw std::runtime_error("Failed to open file: " + path);
    std::vector<double> x, y;
    double a, b;
    while (file >> a >> b) {
        x.push_back(a);
        y.push_back(b);
    }
    m_x = x;
    m_y = y;
}

void FunctionPL::write(const std::string& path) const {
    std::ofstream file{path};
    if (!file) throw std::runtime_error("Failed to open file: " + path);
    file.precision(17);
    for (size_t k = 0; k < m_x.size(); k++) {
        if (k) file << "\n";
        file << m_x[k] << "\t" << m_y[k];
    }
}

void IBI_ForceFunctor::upload() {
    const auto& x = m_force.getXValues();
    const auto& y = m_force.getYValues();
    // the potential counts only on the force's nodes
    const bool pot = m_potential.getXValues() == x && m_potential.getYValues().size() == x.size();
    std::vector<double> now(x);
    now.insert(now.end(), y.begin(), y.end());
    now.push_back(m_force.getLowerDefault());
    now.push_back(m_force.getUpperDefault());
    if (pot) {
        now.insert(now.end(), m_potential.getYValues().begin(), m_potential.getYValues().end());
        now.push_back(m_potential.getLowerDefault());
        now.push_back(m_potential.getUpperDefault());
    }
    if (now == m_sent) return;
    check(p_container.ctx(),
          komd_set_pair_table(p_container.ctx(), (int64_t)x.size(), x.data(), y.data(), m_force.getLowerDefault(),
                              m_force.getUpperDefault(), pot ? m_potential.getYValues().data() : nullptr,
                              m_potential.getLowerDefault(), m_potential.getUpperDefault()),
          "IBI_ForceFunctor");
    m_sent = std::move(now);
}

void IBI_ForceFunctor::handlePairList() {
    upload();
    check(p_container.ctx(), komd_forces_pair(p_container.ctx(), KOMD_PAIR_TABLE), "IBI_ForceFunctor");
}
void Limit::handlePairList() { check(p_container.ctx(), komd_forces_limit(p_container.ctx()), "Limit"); }
void HarmonicBonds::handlePairList() { check(p_container.ctx(), komd_forces_bonded(p_container.ctx()), "HarmonicBonds"); }
void ATM::handleTripleList() { check(p_container.ctx(), komd_forces_3b(p_container.ctx(), m_kind), "ATM"); }

// parameters: a .npy file of KOMD_SUR_NPARAM little-endian float32 values (1-D)
static std::vector<float> read_params(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Could not open surrogate weights: " + path);
    std::vector<char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    size_t off = 0;
    if (buf.size() >= 10 && std::memcmp(buf.data(), "\x93NUMPY", 6) == 0) {
        const int major = (unsigned char)buf[6];
        size_t hl = 0;
        if (major == 1) {
            hl = (unsigned char)buf[8] | ((size_t)(unsigned char)buf[9] << 8);
            off = 10 + hl;
        } else {
            hl = (unsigned char)buf[8] | ((size_t)(unsigned char)buf[9] << 8) | ((size_t)(unsigned char)buf[10] << 16) |
                 ((size_t)(unsigned char)buf[11] << 24);
            off = 12 + hl;
        }
        const std::string header(buf.data() + (major == 1 ? 10 : 12), hl);
        if (header.find("'<f4'") == std::string::npos || header.find("'fortran_order': False") == std::string::npos)
            throw std::runtime_error(path + ": expected a C-ordered float32 array");
    }
    const size_t bytes = buf.size() - off;
    if (bytes != (size_t)KOMD_SUR_NPARAM * 4)
        throw std::runtime_error(path + ": expected " + std::to_string(KOMD_SUR_NPARAM) + " float32 values");
    std::vector<float> p(KOMD_SUR_NPARAM);
    std::memcpy(p.data(), buf.data() + off, bytes);
    return p;
}

ATMSurrogate::ATMSurrogate(MoleculeContainer& container, const std::string& weights, bool hybrid)
    : ForceFunctor3B(container), m_kind(hybrid ? KOMD_3B_HYBRID : KOMD_3B_SURROGATE) {
    const std::vector<float> p = read_params(weights);
    check(container.ctx(), komd_surrogate_set_params(container.ctx(), p.data(), (int64_t)p.size()), "ATMSurrogate");
}
void ATMSurrogate::handleTripleList() {
    check(p_container.ctx(), komd_forces_3b(p_container.ctx(), m_kind), "ATMSurrogate");
}

void Integrator::integrate0() { check(p_container.ctx(), komd_integrate0(p_container.ctx()), "integrate0"); }
void Integrator::integrate1() { check(p_container.ctx(), komd_integrate1(p_container.ctx()), "integrate1"); }

void TemperatureSensor::measure() {
    check(m_container.ctx(), komd_measure_temperature(m_container.ctx(), &m_temperature), "TemperatureSensor");
}

void VelocityScaling::apply() { check(m_container.ctx(), komd_rescale(m_container.ctx(), m_target), "VelocityScaling"); }

static double box_volume(const Configuration& config) {
    return (config.domainHigh[0] - config.domainLow[0]) * (config.domainHigh[1] - config.domainLow[1]) *
           (config.domainHigh[2] - config.domainLow[2]);
}

static std::string sensor_path(const std::string& dir, const std::string& name) {
    return dir.empty() ? name : dir + "/" + name;
}

RDFSensor::RDFSensor(MoleculeContainer& container, const Configuration& config)
    : m_container(container), m_max_r(config.sensor_rdf_max), m_delta_r(config.sensor_rdf_dr),
      m_volume(box_volume(config)), m_dir(config.sensor_dir) {
    const uint64_t num_bins = (uint64_t)(m_max_r / m_delta_r);  // RDFSensor.cpp:19
    m_bins.assign(num_bins, 0.0);
    m_sample.assign(num_bins, 0);
}

void RDFSensor::measure() {
    check(m_container.ctx(),
          komd_rdf(m_container.ctx(), m_max_r, m_delta_r, (int64_t)m_sample.size(), m_sample.data()), "RDFSensor");
    for (size_t k = 0; k < m_bins.size(); k++) m_bins[k] += (double)m_sample[k];
    m_samples += 1;
}

// RDFSensor::write (RDFSensor.cpp:37-57)
void RDFSensor::write(uint64_t simstep) {
    const std::string name = sensor_path(m_dir, "rdf_" + std::to_string(simstep) + ".txt");
    std::ofstream file(name);
    if (!file.is_open()) throw std::runtime_error("Could not write file: " + name);
    const double n = (double)m_container.getNumMolecules();
    const double n_total = n * (n - 1);
    double r = m_delta_r / 2.0;
    const double dr_half = r;
    for (size_t idx = 0; idx < m_bins.size(); idx++) {
        const double n_bin = m_bins[idx] / m_samples;
        const double v_bin = 4.0 / 3.0 * M_PI * (std::pow(r + dr_half, 3) - std::pow(r - dr_half, 3));
        const double g_r = (n_bin * m_volume) / (v_bin * n_total);
        file << r << "\t" << g_r << "\t" << n_bin << "\t" << v_bin << "\t" << n_total << "\t" << m_volume << "\t"
             << (n_bin * m_volume) << "\t" << (v_bin * n_total) << "\n";
        r += m_delta_r;
    }
}

PressureSensor::PressureSensor(MoleculeContainer& container, const Configuration& config, int kind3b)
    : m_container(container), m_kind3b(kind3b), m_volume(box_volume(config)), m_dir(config.sensor_dir) {}

void PressureSensor::measure() {
    komd_observables o{};
    check(m_container.ctx(), komd_measure_observables(m_container.ctx(), m_kind3b, &o), "PressureSensor");
    for (int k = 0; k < 6; k++) m_pressures[(size_t)k] = o.kinetic_tensor[k] + o.virial_tensor[k];
}

// PressureSensor::write (PressureSensor.cpp:34-70): tensor / volume in Pa,
// factor = conv_Da_kg / (ps_s^2 A_m) = 1.66053906892e-27 / 1e-34.
void PressureSensor::write(uint64_t simstep) {
    const std::string name = sensor_path(m_dir, "pressures_" + std::to_string(simstep) + ".txt");
    std::ofstream file(name);
    if (!file.is_open()) throw std::runtime_error("Could not write file: " + name);
    const double factor = 1.66053906892 * std::pow(10.0, 7.0);
    int next_count = 1, count = 0;
    for (int idx = 0; idx < 6; idx++) {
        file << m_pressures[(size_t)idx] / m_volume * factor;
        count++;
        if (count == next_count) {
            count = 0;
            next_count++;
            file << std::endl;
        } else {
            file << " ";
        }
    }
}

std::unique_ptr<Registry> makeRegistry(const Configuration& config) {
    auto reg = std::make_unique<Registry>();
    reg->config = config;
    auto lc = std::make_unique<LinkedCells>(config);
    CheckpointIO::loadCheckpoint(config, lc->sites());
    lc->init();
    reg->container = std::move(lc);
    MoleculeContainer& c = *reg->container;
    // the pair functor (main.cpp:56-57): LJ 12-6, ATM2B with enable_3b_approx, or the
    // tabulated force of a pair_table file (the IBI functor the plugin registers in LJ's
    // place, IBI.cpp:51-55, with the plugin's defaults 1e6 below and 0 above the nodes)
    if (!config.pair_table.empty()) {
        auto ibi = std::make_shared<IBI_ForceFunctor>(c);
        ibi->getForceFunction().read(config.pair_table);
        if (!config.pair_table_potential.empty()) ibi->getPotentialFunction().read(config.pair_table_potential);
        ibi->upload();
        reg->forceFunctors.push_back(ibi);
        check(c.ctx(), komd_select_pair(c.ctx(), KOMD_PAIR_TABLE), "pair_table");
    } else if (config.enable_3b_approx) {
        reg->forceFunctors.push_back(std::make_shared<ATM2B>(c));
        check(c.ctx(), komd_select_pair(c.ctx(), KOMD_PAIR_ATM2B), "enable_3b_approx");
    } else {
        reg->forceFunctors.push_back(std::make_shared<LJ12_6>(c));
    }
    reg->forceFunctors.push_back(std::make_shared<Limit>(c));
    if (config.enable_3b) {
        c.enable3B();
        if (!config.surrogate_weights.empty()) {
            const bool hybrid = config.hybrid_axis >= 0;
            if (hybrid)
                check(c.ctx(), komd_set_hybrid(c.ctx(), config.hybrid_axis, config.hybrid_center, config.hybrid_half,
                                               config.hybrid_blend),
                      "hybrid_region");
            reg->forceFunctors3b.push_back(std::make_shared<ATMSurrogate>(c, config.surrogate_weights, hybrid));
            if (config.judge_every > 0)
                check(c.ctx(), komd_set_judge(c.ctx(), config.judge_every, config.judge_max), "judge_every");
        } else {
            reg->forceFunctors3b.push_back(std::make_shared<ATM>(c, config.atm_mode == "reference"));
        }
    }
    reg->integrators.push_back(std::make_unique<Integrator>(c));
    reg->temperatureSensor = std::make_unique<TemperatureSensor>(c);
    // sensors in the order main.cpp:81-84 registers them
    if (config.enable_sensor_rdf) reg->sensors.push_back(std::make_shared<RDFSensor>(c, config));
    if (config.enable_sensor_pres) {
        int kind = 0;
        if (c.threeBody() && !reg->forceFunctors3b.empty()) {
            if (auto a = std::dynamic_pointer_cast<ATM>(reg->forceFunctors3b[0])) kind = a->kind();
            if (auto s = std::dynamic_pointer_cast<ATMSurrogate>(reg->forceFunctors3b[0])) kind = s->kind();
        }
        reg->sensors.push_back(std::make_shared<PressureSensor>(c, config, kind));
    }
    if (config.thermostat) reg->thermostats.push_back(std::make_unique<VelocityScaling>(c, config.temperature));
    return reg;
}

void Simulation::run(Registry& reg, uint64_t steps) {
    MoleculeContainer& c = *reg.container;
    // initial pass to compute forces
    c.updateContainer();
    for (auto& p : reg.forceFunctors) (*p)();
    if (c.threeBody())
        for (auto& p : reg.forceFunctors3b) (*p)();
    reg.temperatureSensor->measure();
    for (auto& t : reg.thermostats) t->apply();
    // main loop
    for (uint64_t s = 0; s < steps; s++) {
        for (auto& i : reg.integrators) i->integrate0();
        c.updateContainer();
        for (auto& p : reg.forceFunctors) (*p)();
        if (c.threeBody())
            for (auto& p : reg.forceFunctors3b) (*p)();
        for (auto& i : reg.integrators) i->integrate1();
        for (auto& sensor : reg.sensors) sensor->measure();
        reg.temperatureSensor->measure();
        m_simstep++;
        if (reg.config.write_freq > 0 && m_simstep % reg.config.write_freq == 0)
            for (auto& sensor : reg.sensors) sensor->write(m_simstep);
        for (auto& t : reg.thermostats) t->apply();
    }
    for (auto& sensor : reg.sensors) sensor->write(m_simstep);
}

void Simulation::runDevice(Registry& reg, uint64_t steps) {
    MoleculeContainer& c = *reg.container;
    if (!reg.sensors.empty())
        throw std::runtime_error("Simulation::runDevice: sensors measure every step; use the per-functor loop");
    int kind = 0;
    if (c.threeBody() && !reg.forceFunctors3b.empty()) {
        if (auto a = std::dynamic_pointer_cast<ATM>(reg.forceFunctors3b[0])) kind = a->kind();
        if (auto s = std::dynamic_pointer_cast<ATMSurrogate>(reg.forceFunctors3b[0])) kind = s->kind();
    }
    const double t0 = reg.thermostats.empty() ? 0.0 : reg.config.temperature;
    int pair = KOMD_PAIR_LJ12_6;
    if (!reg.forceFunctors.empty()) {
        if (auto ibi = std::dynamic_pointer_cast<IBI_ForceFunctor>(reg.forceFunctors[0])) {
            ibi->upload();
            pair = KOMD_PAIR_TABLE;
        } else if (std::dynamic_pointer_cast<ATM2B>(reg.forceFunctors[0])) {
            pair = KOMD_PAIR_ATM2B;
        }
    }
    check(c.ctx(), komd_select_pair(c.ctx(), pair), "Simulation::runDevice");
    // the reference reuses its pair list for pair_steps steps without a skin; the device loop
    // rebuilds every step unless the deck gives a skin (Verlet lists, no pair missed)
    check(c.ctx(), komd_set_skin(c.ctx(), reg.config.skin), "Simulation::runDevice");
    check(c.ctx(), komd_compute_forces(c.ctx(), kind, 0), "Simulation::runDevice");
    if (t0 > 0) {
        reg.temperatureSensor->measure();
        for (auto& t : reg.thermostats) t->apply();
    }
    check(c.ctx(), komd_run(c.ctx(), (int64_t)steps, kind, 0, t0), "Simulation::runDevice");
    m_simstep += steps;
}

}  // namespace komd_host
