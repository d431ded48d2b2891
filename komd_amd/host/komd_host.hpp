// C++ host layer above libkomd_hip.so: the simulation objects of a KoMD run (container,
// force functors, integrator, temperature sensor, thermostat, simulation loop, deck and
// checkpoint I/O) with the reference's names and call order, each one calling the C ABI
// (include/komd.h).  Errors surface as std::runtime_error carrying komd_last_error(),
// as the reference reports failures by exceptions and error logs.
//
// Reference interfaces mirrored (KoMD src/):
//   ForceFunctor / ForceFunctor3B   potentials/ForceFunctor.h:15-33, ForceFunctor3B.h:14-34
//   MoleculeContainer / LinkedCells  container/MoleculeContainer.h:19-146, LinkedCells.cpp:11-81
//   Integrator                       integrators/Integrator.h:14-65
//   TemperatureSensor                sensors/TemperatureSensor.cpp:22-43
//   Sensor / RDFSensor / PressureSensor  sensors/Sensor.h, RDFSensor.cpp:13-57, PressureSensor.cpp:21-70
//   Thermostat / VelocityScaling     thermostats/VelocityScaling.cpp:12-25
//   Simulation::run                  Simulation.cpp:10-80
//   ATM2B                            potentials/ATM2B.cpp:9-56
//   FunctionPL / IBI_ForceFunctor    plugins/IBI/IBI_Math.h:13-78, IBI_Math.cpp, IBI_ForceFunctor.cpp:10-42
//   FileInput::readFile              IO/FileInput.cpp:11-121
//   CheckpointIO                     IO/CheckpointIO.cpp:13-76
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/komd.h"

namespace komd_host {

using d3 = std::array<double, 3>;

// Unit conversions of the reference (util/constants.h:12-20, TemperatureSensor.cpp:39-42):
// epsilon and energy_3b are read in J and J nm^9; the internal energy unit is u A^2/ps^2.
constexpr double kConvJEi = 6.022140754e22;
// T = factor * sum(m v.v) / (3 N), factor = (A/ps in m/s)^2 * (Da in kg) / k_B
constexpr double kTemperatureFactor = 1e4 * (1.66053906892e-27 / 1.380649e-23);

// Deck keys used by the force path (Configuration.h, defaults.h), plus the keys of the
// GPU path's own three-body options.
struct Configuration {
    double temperature = 300;
    double delta_t = 0.001;
    uint64_t timesteps = 1;
    double cutoff = 2.5 * 3.41;
    double cell_size = 3.0 * 3.41;
    double limit_factor = 20;
    double energy_3b = 1;
    d3 domainLow{0, 0, 0};
    d3 domainHigh{10, 10, 10};
    uint64_t write_freq = 10;
    bool storeCheckpoint = false;
    bool enable_3b = false;
    bool enable_3b_direct = true;
    bool enable_3b_approx = false;       // ATM2B in LJ's functor slot (main.cpp:56-57)
    std::vector<std::pair<std::string, d3>> checkpoint_files;
    // GPU path options (no reference counterpart)
    std::string atm_mode = "reference";  // "reference": ATM_NOLIST semantics; "exact": physical, all triplets
    std::string surrogate_weights;       // three-body surrogate parameters (.f32, KOMD_SUR_NPARAM floats)
    int pair_steps = 4;                  // reference: pair list reused for this many steps (PairList.h:22-24)
    int triple_steps = 8;                // reference: triple list likewise
    double skin = 0.0;                   // device loop: Verlet-list skin (0: lists every step), komd_set_skin
    int64_t judge_every = 0;             // exact-ATM judge interval of surrogate runs
    double judge_max = 0.05;             // fallback threshold (relative RMS force error)
    int hybrid_axis = -1;                // >= 0: hybrid exact/surrogate region along this axis
    double hybrid_center = 0, hybrid_half = 0, hybrid_blend = 2.0;
    bool thermostat = true;              // velocity rescaling every step (main.cpp:93)
    int checkpoint_precision = 17;       // digits of the .ps writer
    std::string pair_table;              // tabulated pair force (FunctionPL file) in LJ's slot
    std::string pair_table_potential;    // its potential on the same nodes (energies only)
    // sensors (defaults.h:30-46)
    bool enable_sensor_rdf = false;
    bool enable_sensor_pres = false;
    double sensor_rdf_max = 5.0;
    double sensor_rdf_dr = 0.1;
    std::string sensor_dir;              // directory of the sensor files (reference: the working directory)
};

struct Component {  // one COMP line per site: eps (J), sigma, mass, offset
    struct Site {
        double eps, sig, mass;
        d3 r;
    };
    std::vector<Site> sites;
};

// FileInput::readFile: false (with the message) on an unreadable deck or a failed check.
bool readFile(const std::string& path, Configuration& config, std::vector<Component>& components, std::string* err);

// Host copy of the sites in the reference's layout (SOA.h:11-38): interleaved d3 r and v,
// per-site mass / epsilon (internal units) / sigma, molecule id per site.
struct Sites {
    std::vector<double> r, v, f, mass, eps, sig;
    std::vector<uint64_t> id;
    size_t size() const { return mass.size(); }
    void add(uint64_t mol_id, const d3& pos, const d3& vel, double m, double e, double s);
};

class MoleculeContainer;

namespace CheckpointIO {
// loadCheckpoint: every checkpoint_file of the deck, positions shifted by its offset.
void loadCheckpoint(const Configuration& config, Sites& sites);
void readPs(const std::string& path, const d3& offset, Sites& sites);
// writeCheckpoint: "ckpt_<simstep>.ps" (or `path`), one site per line, `precision` digits.
void writeCheckpoint(uint64_t simstep, MoleculeContainer& container, int precision, const std::string& path = "");
void writePs(const std::string& path, const Sites& sites, int precision);
}  // namespace CheckpointIO

// MoleculeContainer: owns the device context.
class MoleculeContainer {
public:
    virtual ~MoleculeContainer();
    virtual void init() = 0;
    virtual void updateContainer() = 0;
    uint64_t getNumMolecules() const { return m_num_molecules; }
    // constructSOAs: host sites -> device; writeSOA2AOS: device -> host sites
    void constructSOAs();
    void writeSOA2AOS();
    Sites& sites() { return m_sites; }
    komd_ctx* ctx() const { return m_ctx; }
    void enable3B() { m_enable_3b = true; }
    void disable3B() { m_enable_3b = false; }
    bool threeBody() const { return m_enable_3b; }

protected:
    Sites m_sites;
    komd_ctx* m_ctx = nullptr;
    uint64_t m_num_molecules = 0;
    bool m_enable_3b = false;
};

class LinkedCells : public MoleculeContainer {
public:
    explicit LinkedCells(const Configuration& config);
    void init() override;
    void updateContainer() override;
    std::array<int64_t, 3> gridDims() const;

private:
    komd_box m_box{};
    komd_params m_params{};
};

// ForceFunctor: operator() runs the pass and waits for it (ForceFunctor.cpp:10-14).
class ForceFunctor {
public:
    explicit ForceFunctor(MoleculeContainer& container) : p_container(container) {}
    virtual ~ForceFunctor() = default;
    void operator()() { handlePairList(); }

protected:
    virtual void handlePairList() = 0;
    MoleculeContainer& p_container;
};

class LJ12_6 : public ForceFunctor {
public:
    using ForceFunctor::ForceFunctor;

protected:
    void handlePairList() override;
};

// ATM2B (potentials/ATM2B.cpp:9-56): LJ 12-6 times 1 - 2 nu rho / (3 eps sig^6), registered
// in LJ's place when the deck sets enable_3b_approx (main.cpp:56-57).
class ATM2B : public ForceFunctor {
public:
    using ForceFunctor::ForceFunctor;

protected:
    void handlePairList() override;
};

// FunctionPL (plugins/IBI/IBI_Math.h:13-78, IBI_Math.cpp): piecewise-linear function of
// ascending nodes x with values y, the lower default below the first node and the upper
// default above the last; read / write use the reference's two-column text format.
class FunctionPL {
public:
    explicit FunctionPL(size_t buffer_size = 0, double def_low = 0.0, double def_high = 0.0)
        : m_x(buffer_size, 0.0), m_y(buffer_size, 0.0), m_low(def_low), m_high(def_high) {}
    void setXValues(const std::vector<double>& x) { m_x = x; }
    void setYValues(const std::vector<double>& y) { m_y = y; }
    std::vector<double>& getXValues() { return m_x; }
    std::vector<double>& getYValues() { return m_y; }
    const std::vector<double>& getXValues() const { return m_x; }
    const std::vector<double>& getYValues() const { return m_y; }
    double getLowerDefault() const { return m_low; }
    double getUpperDefault() const { return m_high; }
    void setLowerDefault(double v) { m_low = v; }
    void setUpperDefault(double v) { m_high = v; }
    // evaluateAt (IBI_Math.h:30-41); at x equal to the first node: the first value
    double evaluateAt(double x) const;
    // forward / backward differences at the ends, central ones inside (IBI_Math.cpp:22-35)
    void derivative(FunctionPL& target) const;
    void read(const std::string& path);
    void write(const std::string& path) const;

private:
    std::vector<double> m_x, m_y;
    double m_low, m_high;
};

// IBI_ForceFunctor<IBI_Default> (plugins/IBI/IBI_ForceFunctor.cpp:10-42): the tabulated pair
// force.  Its tables may be edited in place between steps (as the IBI plugin does), so each
// call copies them to the device when they differ from the last copy.
class IBI_ForceFunctor : public ForceFunctor {
public:
    explicit IBI_ForceFunctor(MoleculeContainer& container)
        : ForceFunctor(container), m_force(0, 1e+6, 0.0), m_potential(0, 1e+6, 0.0) {}
    FunctionPL& getForceFunction() { return m_force; }
    FunctionPL& getPotentialFunction() { return m_potential; }
    void upload();

protected:
    void handlePairList() override;

private:
    FunctionPL m_force, m_potential;
    std::vector<double> m_sent;  // the tables and defaults of the last copy
};

class Limit : public ForceFunctor {
public:
    using ForceFunctor::ForceFunctor;

protected:
    void handlePairList() override;
};

class HarmonicBonds : public ForceFunctor {  // new physics in FENE's functor slot
public:
    using ForceFunctor::ForceFunctor;

protected:
    void handlePairList() override;
};

class ForceFunctor3B {
public:
    explicit ForceFunctor3B(MoleculeContainer& container) : p_container(container) {}
    virtual ~ForceFunctor3B() = default;
    void operator()() { handleTripleList(); }

protected:
    virtual void handleTripleList() = 0;
    MoleculeContainer& p_container;
};

// ATM_NOLIST (reference semantics) and ATM (physical, complete) share one functor.
class ATM : public ForceFunctor3B {
public:
    ATM(MoleculeContainer& container, bool reference_semantics)
        : ForceFunctor3B(container), m_kind(reference_semantics ? KOMD_3B_ATM_REF : KOMD_3B_ATM_EXACT) {}
    int kind() const { return m_kind; }

protected:
    void handleTripleList() override;
    int m_kind;
};

class ATMSurrogate : public ForceFunctor3B {
public:
    ATMSurrogate(MoleculeContainer& container, const std::string& weights, bool hybrid);
    int kind() const { return m_kind; }

protected:
    void handleTripleList() override;
    int m_kind;
};

class Integrator {
public:
    explicit Integrator(MoleculeContainer& container) : p_container(container) {}
    virtual ~Integrator() = default;
    virtual void integrate0();
    virtual void integrate1();

protected:
    MoleculeContainer& p_container;
};

class TemperatureSensor {
public:
    explicit TemperatureSensor(MoleculeContainer& container) : m_container(container) {}
    void measure();
    double getTemperature() const { return m_temperature; }

private:
    MoleculeContainer& m_container;
    double m_temperature = 0;
};

// Sensor (sensors/Sensor.h): measure() after integrate1 every step, write(simstep) every
// write_freq steps and at the end (Simulation.cpp:58-79).
class Sensor {
public:
    virtual ~Sensor() = default;
    virtual void measure() = 0;
    virtual void write(uint64_t simstep) = 0;
};

// RDFSensor (RDFSensor.cpp:13-57): ordered-pair counts of molecule COM distances summed
// over samples; write() prints "rdf_<simstep>.txt" with the reference's eight columns.
class RDFSensor : public Sensor {
public:
    RDFSensor(MoleculeContainer& container, const Configuration& config);
    void measure() override;
    void write(uint64_t simstep) override;
    const std::vector<double>& bins() const { return m_bins; }
    double samples() const { return m_samples; }

private:
    MoleculeContainer& m_container;
    double m_max_r, m_delta_r, m_volume;
    std::string m_dir;
    std::vector<double> m_bins;
    std::vector<uint64_t> m_sample;
    double m_samples = 0;
};

// PressureSensor (PressureSensor.cpp:21-70): the stress tensor sum m v v + virial of the
// registered potentials, XX XY YY XZ YZ ZZ; write() prints "pressures_<simstep>.txt" in Pa,
// as a lower triangle (one, two, three values per line).
class PressureSensor : public Sensor {
public:
    PressureSensor(MoleculeContainer& container, const Configuration& config, int kind3b);
    void measure() override;
    void write(uint64_t simstep) override;
    const std::array<double, 6>& pressures() const { return m_pressures; }

private:
    MoleculeContainer& m_container;
    int m_kind3b;
    double m_volume;
    std::string m_dir;
    std::array<double, 6> m_pressures{};
};

class Thermostat {
public:
    virtual ~Thermostat() = default;
    virtual void apply() = 0;
};

class VelocityScaling : public Thermostat {
public:
    VelocityScaling(MoleculeContainer& container, double target) : m_container(container), m_target(target) {}
    void apply() override;

private:
    MoleculeContainer& m_container;
    double m_target;
};

// The objects of one run, filled in the order main.cpp:48-93 registers them.
struct Registry {
    Configuration config;
    std::unique_ptr<MoleculeContainer> container;
    std::vector<std::shared_ptr<ForceFunctor>> forceFunctors;
    std::vector<std::shared_ptr<ForceFunctor3B>> forceFunctors3b;
    std::vector<std::unique_ptr<Integrator>> integrators;
    std::vector<std::unique_ptr<Thermostat>> thermostats;
    std::unique_ptr<TemperatureSensor> temperatureSensor;
    std::vector<std::shared_ptr<Sensor>> sensors;
};

// Builds the registry of a deck: container from the checkpoint files, LJ (+ harmonic
// bonds when a topology is set later) + Limit, the three-body functor of the deck, the
// integrator, temperature sensor and (optionally) the rescaling thermostat.
std::unique_ptr<Registry> makeRegistry(const Configuration& config);

class Simulation {
public:
    // The reference loop (Simulation.cpp:49-74), one functor call after the other.
    void run(Registry& reg, uint64_t steps);
    // The same loop queued on the device as one replayed step (komd_run).  Sensors
    // measure between functors, so a registry with sensors must use run().
    void runDevice(Registry& reg, uint64_t steps);
    uint64_t simstep() const { return m_simstep; }

private:
    uint64_t m_simstep = 0;
};

// status check of a C-ABI call
void check(komd_ctx* ctx, komd_status st, const char* what);

}  // namespace komd_host
