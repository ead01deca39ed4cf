// Implementation of the reference-shaped host API (lgf_host.hpp) over the
// C-ABI of include/lgf_flow.h.  File names, directory layout, CSV and q(x)
// formats follow the reference byte for byte (see the citations below).
#include "lgf_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>

namespace fs = std::filesystem;

namespace lgf_host {

static lgf_ctx* g_ctx = nullptr;
static int g_device = 0;
static int g_precision = LGF_FP64;
static int g_measure_flags = LGF_MEASURE_FP64;
static std::string g_input_method = "LHMC_3_of_0_330000";  // LGF/LGF.cpp:164

void set_device(int device) { g_device = device; }
void set_default_precision(int p) { g_precision = p; }
int default_precision() { return g_precision; }
void set_measure_flags(int f) { g_measure_flags = f; }
int measure_flags() { return g_measure_flags; }
void set_input_method(const std::string& m) { g_input_method = m; }
std::string input_method() { return g_input_method; }
static int g_checkpoint_every = 0;
void set_checkpoint_every(int n) {
    if (n < 0) throw Error(LGF_E_ARG, "checkpoint interval must be >= 0");
    g_checkpoint_every = n;
}
int checkpoint_every() { return g_checkpoint_every; }
static int g_autotune_steps = 0;
void set_autotune(int nsteps) {
    if (nsteps < 0) throw Error(LGF_E_ARG, "autotune step count must be >= 0");
    g_autotune_steps = nsteps;
}
int autotune_steps() { return g_autotune_steps; }

void check(int rc, const char* what) {
    if (rc != LGF_OK) {
        std::string msg = std::string(what) + " failed (status " + std::to_string(rc) + ")";
        if (g_ctx) msg += ": " + std::string(lgf_ctx_last_error(g_ctx));
        throw Error(rc, msg);
    }
}

lgf_ctx* context() {
    if (!g_ctx) {
        int rc = lgf_ctx_create(g_device, &g_ctx);
        if (rc != LGF_OK) throw Error(rc, "no usable GPU device " + std::to_string(g_device));
    }
    return g_ctx;
}

// std::to_string(beta) with '.' -> '_' (LGF/Field/SU3_field.cpp:262-263)
std::string beta_string(double beta) {
    std::string b = std::to_string(beta);
    std::replace(b.begin(), b.end(), '.', '_');
    return b;
}

static std::string shape_dir(const int* shape) {
    std::string s;
    for (int i = 0; i < 3; ++i) s += std::to_string(shape[i]) + "X";
    return s + std::to_string(shape[3]);
}

// LGF/Field/SU3_field.cpp:270-284
std::string config_path(double beta, const int* shape, const std::string& method, const std::string& folder,
                        const std::string& id) {
    return folder + "ensembles" + "/su3/beta" + beta_string(beta) + "/" + shape_dir(shape) + "/" + method + "/" +
           "torus_extdof4_" + id + ".bin";
}

static std::string observable_dir(const std::string& kind, double beta, const int* shape, const std::string& method,
                                  const std::string& folder) {
    return folder + "Observables" + "/" + kind + "/beta" + beta_string(beta) + "/" + shape_dir(shape) + "/" + method +
           "/";
}

}  // namespace lgf_host

using namespace lgf_host;

// ---------------------------------------------------------------------------
SimulationParameters::SimulationParameters(int NrDims, int extdofs, std::vector<int> shape, double beta,
                                           std::vector<int> cuts, std::string LatticeType)
    : m_NrDims(NrDims), m_extdofs(extdofs), m_shape(shape), m_beta(beta), m_cuts(cuts), m_LatticeType(LatticeType) {
    if (m_LatticeType != "torus" && m_LatticeType != "torus_OpenT")
        throw Error(LGF_E_ARG, "Lattice type was not recognised! Try <torus> or <torus_OpenT>");
    if (m_LatticeType == "torus_OpenT")
        throw Error(LGF_E_UNSUPPORTED, "lattice type torus_OpenT is not supported by the GPU flow path");
    if (m_NrDims != 4 || (int)m_shape.size() != 4) throw Error(LGF_E_SHAPE, "the flow path needs NrDims = 4");
    if (m_extdofs != 4) throw Error(LGF_E_ARG, "the gauge field needs extdofs = 4");
}

Lattice::Lattice(SimulationParameters& params, bool) : m_type(params.getLatticeType()) {
    m_totalVolume = 1;
    for (int i = 0; i < 4; ++i) {
        m_shape[i] = params.getshape(i);
        if (m_shape[i] <= 0 || m_shape[i] % 2 != 0)  // LGF/Lattice/Lattice.cpp:42-49
            throw Error(LGF_E_SHAPE, "Even- Odd scheme requires an even number of lattice points in each direction "
                                     "with the torus topology!");
        m_totalVolume *= m_shape[i];
    }
}

// ---------------------------------------------------------------------------
namespace {
// per-field measurement cache: one clover sweep serves Q and E at one flow time
struct MeasCache {
    uint64_t version = ~0ull;
    double obs[LGF_NUM_OBS] = {0, 0, 0, 0, 0};
};
std::map<const SU3_field*, uint64_t> g_version;
std::map<const SU3_field*, MeasCache> g_cache;
}  // namespace

static void bump(const SU3_field* f) { g_version[f]++; }

SU3_field::SU3_field(Lattice& lattice, int NrExtDOF) : m_lattice(&lattice), m_precision(default_precision()) {
    if (NrExtDOF != 4) throw Error(LGF_E_ARG, "SU3_field needs NrExtDOF = 4");
    check(lgf_field_create(context(), lattice.getShape(), m_precision, &m_field), "lgf_field_create");
    g_version[this] = 0;
}

SU3_field::~SU3_field() {
    if (m_field) lgf_field_destroy(m_field);
    g_version.erase(this);
    g_cache.erase(this);
}

void SU3_field::loadSU3FromFile(double beta, const std::string& updateMethod, const std::string& dataFolder,
                                const std::string& identifier) {
    std::string path = config_path(beta, m_lattice->getShape(), updateMethod, dataFolder, identifier);
    fs::create_directories(fs::path(path).parent_path());  // the reference creates them on load too
    std::cout << "Loading SU3 Field onto 1 processes\nFrom path " << path << "\n";
    std::cout.flush();
    check(lgf_field_load_file(m_field, path.c_str()), "loadSU3FromFile");
    std::cout << "File loaded \n";
    std::cout.flush();
    bump(this);
}

void SU3_field::saveSU3ToFile(double beta, const std::string& updateMethod, const std::string& dataFolder,
                              const std::string& identifier) {
    std::string path = config_path(beta, m_lattice->getShape(), updateMethod, dataFolder, identifier);
    fs::create_directories(fs::path(path).parent_path());
    std::cout << "Saving SU3 Field to path " << path << "\n";
    std::cout.flush();
    check(lgf_field_save_file(m_field, path.c_str()), "saveSU3ToFile");
}

void SU3_field::InitializeColdStart() {
    check(lgf_field_set_cold(m_field), "InitializeColdStart");
    bump(this);
}

double SU3_field::Avg_Plaquette() {
    double p, umax, umean;
    check(lgf_diagnostics(m_field, &p, &umax, &umean), "Avg_Plaquette");
    return p;
}

// ---------------------------------------------------------------------------
GradientFlow::GradientFlow(Action&, SU3_field& U, double epsilon, int FlowStepsPerMeasurement)
    : m_U(&U), m_FlowStepsPerMeasurement(FlowStepsPerMeasurement), m_epsilon(epsilon) {}

// LGF/Gradient_Flow/GradientFlow.cpp:22-60
void GradientFlow::flow() {
    check(lgf_flow_steps(m_U->handle(), m_epsilon, 1), "GradientFlow::flow");
    bump(m_U);
    m_flowTime += m_epsilon;
    m_NrSteps++;
    if (m_NrSteps % m_FlowStepsPerMeasurement == 0) MakeMeasurements();
}

void GradientFlow::MakeMeasurements() {
    if (m_topologicalCharge) m_topologicalCharge->calculate(m_flowTime);
    if (m_EnergyDensity) m_EnergyDensity->calculate(m_flowTime);
}

void FieldMeasurement::observables(SU3_field& U, double* obs) {
    MeasCache& c = g_cache[&U];
    const uint64_t v = g_version[&U];
    if (c.version != v) {
        int rc = lgf_measure_observables(U.handle(), measure_flags(), c.obs, nullptr);
        if (rc != LGF_OK && rc != LGF_E_NAN) check(rc, "measure");  // NaN is printed like the reference does
        c.version = v;
    }
    for (int k = 0; k < LGF_NUM_OBS; ++k) obs[k] = c.obs[k];
}

void FieldMeasurement::measure(SU3_field& U, double, double* E, double* Q, double* Qabs) {
    double o[LGF_NUM_OBS];
    observables(U, o);
    *E = o[LGF_OBS_E];
    *Q = o[LGF_OBS_Q];
    *Qabs = o[LGF_OBS_QABS];
}

// ---------------------------------------------------------------------------
// CSV "t,value" lines with default stream formatting, appended
// (LGF/Observables/EnergyDensity.cpp:37-80, TopologicalCharge.cpp:31-71),
// plus a full-precision twin next to it.
static void write_csv(const std::string& dir, const std::string& file, const std::vector<double>& t,
                      const std::vector<double>& v) {
    fs::create_directories(dir);
    {
        std::ofstream out(dir + file + ".csv", std::ios_base::app);
        if (!out) throw Error(LGF_E_IO, "cannot write " + dir + file + ".csv");
        for (size_t i = 0; i < v.size(); ++i) out << t[i] << "," << v[i] << "\n";
    }
    std::ofstream full(dir + file + ".full.csv", std::ios_base::app);
    if (!full) throw Error(LGF_E_IO, "cannot write " + dir + file + ".full.csv");
    full.precision(17);
    for (size_t i = 0; i < v.size(); ++i) full << t[i] << "," << v[i] << "\n";
}

void EnergyDensity::record(double flowTime, double value) {
    m_resultVector.push_back(value);
    m_FlowMeasurementTimeVector.push_back(flowTime);
    std::cout << "Energy Density\t   = " << value << "\n";  // LGF/Observables/EnergyDensity.cpp:30
    std::cout.flush();
}

void EnergyDensity::calculate(double flowTime) {
    double E, Q, Qa;
    FieldMeasurement::measure(*m_U, flowTime, &E, &Q, &Qa);
    record(flowTime, E);
}

void EnergyDensity::saveEnergyDensityToFile(double beta, const std::string& updateMethod,
                                            const std::string& dataFolder, const std::string& identifier) {
    write_csv(observable_dir("EnergyDensity", beta, m_U->getLatticePtr().getShape(), updateMethod, dataFolder),
              "torus_extdof4" + identifier, m_FlowMeasurementTimeVector, m_resultVector);
}

void TopologicalCharge::record(double flowTime, double value) {
    m_resultVector.push_back(value);
    m_FlowMeasurementTimeVector.push_back(flowTime);
    std::cout << "Topological Charge = " << value << "\n";  // LGF/Observables/TopologicalCharge.cpp:25
    std::cout.flush();
}

void TopologicalCharge::calculate(double flowTime) {
    double E, Q, Qa;
    FieldMeasurement::measure(*m_U, flowTime, &E, &Q, &Qa);
    record(flowTime, Q);
}

void TopologicalCharge::saveTopologicalChargeToFile(double beta, const std::string& updateMethod,
                                                    const std::string& dataFolder, const std::string& identifier) {
    write_csv(observable_dir("Topological_Charge", beta, m_U->getLatticePtr().getShape(), updateMethod, dataFolder),
              "torus_extdof4" + identifier, m_FlowMeasurementTimeVector, m_resultVector);
}

// LGF/Observables/ImprovedEnergyDensity.cpp:9-74: prints like EnergyDensity,
// files under Observables/ImprovedEnergyDensity/
void ImprovedEnergyDensity::calculate(double flowTime) {
    double o[LGF_NUM_OBS];
    FieldMeasurement::observables(*m_U, o);
    m_resultVector.push_back(o[LGF_OBS_E_IMPROVED]);
    m_FlowMeasurementTimeVector.push_back(flowTime);
    std::cout << "Energy Density\t   = " << o[LGF_OBS_E_IMPROVED] << "\n";  // ImprovedEnergyDensity.cpp:29
    std::cout.flush();
}

void ImprovedEnergyDensity::saveEnergyDensityToFile(double beta, const std::string& updateMethod,
                                                    const std::string& dataFolder, const std::string& identifier) {
    write_csv(observable_dir("ImprovedEnergyDensity", beta, m_U->getLatticePtr().getShape(), updateMethod,
                             dataFolder),
              "torus_extdof4" + identifier, m_FlowMeasurementTimeVector, m_resultVector);
}

// LGF/Observables/UnImprovedTopologicalCharge.cpp:8-67: the charge with the
// 1x1 clover F, files under Observables/UnImprovedTopological_Charge/
void UnImprovedTopologicalCharge::calculate(double flowTime) {
    double o[LGF_NUM_OBS];
    FieldMeasurement::observables(*m_U, o);
    m_resultVector.push_back(o[LGF_OBS_Q_UNIMPROVED]);
    m_FlowMeasurementTimeVector.push_back(flowTime);
    std::cout << "Topological Charge = " << o[LGF_OBS_Q_UNIMPROVED] << "\n";  // UnImprovedTopologicalCharge.cpp:25
    std::cout.flush();
}

void UnImprovedTopologicalCharge::saveTopologicalChargeToFile(double beta, const std::string& updateMethod,
                                                              const std::string& dataFolder,
                                                              const std::string& identifier) {
    write_csv(observable_dir("UnImprovedTopological_Charge", beta, m_U->getLatticePtr().getShape(), updateMethod,
                             dataFolder),
              "torus_extdof4" + identifier, m_FlowMeasurementTimeVector, m_resultVector);
}

// ---------------------------------------------------------------------------
QdensityDistribution::QdensityDistribution(SU3_field& U, double beta, const std::string& updateMethod,
                                           const std::string& dataFolder, const std::string& identifier)
    : m_U(&U), m_beta(beta), m_updateMethod(updateMethod), m_dataFolder(dataFolder), m_identifier(identifier) {}

// q(x) at the global index, scaled by (1.0/4.0*PI2) = pi^2/4 exactly as the
// reference's file (LGF/Observables/QdensityDistribution.cpp:24 and
// LGF/Field/Double_field.cpp:17-98).
void QdensityDistribution::calculate(double) {
    static const double PI = 4.0 * atan(1.0);
    static const double PI2 = PI * PI;
    const long long V = m_U->getLatticePtr().m_totalVolume;
    std::vector<double> q((size_t)V);
    double E, Q, Qa;
    int rc = lgf_measure(m_U->handle(), measure_flags(), &E, &Q, &Qa, q.data());
    if (rc != LGF_OK && rc != LGF_E_NAN) check(rc, "QdensityDistribution");
    for (auto& x : q) x = (1.0 / 4.0 * PI2) * x;
    std::string dir = observable_dir("Qdensity", m_beta, m_U->getLatticePtr().getShape(), m_updateMethod, m_dataFolder);
    fs::create_directories(dir);
    std::string path = dir + "torus_extdof1_" + m_identifier + ".bin";
    std::cout << "Saving Double Field to path " << path << "\n";
    std::cout.flush();
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    out.write(reinterpret_cast<const char*>(q.data()), (std::streamsize)(q.size() * sizeof(double)));
    if (!out) throw Error(LGF_E_IO, "cannot write " + path);
}

// Tunes the context for `n` lattices of the run's shape in the flow
// precision, once per (shape, n, precision), before the fields of a driver
// exist (a batch takes the tile order in force when it is created).
static void tune_for(SimulationParameters& params, int n) {
    if (lgf_host::autotune_steps() <= 0 || n <= 0) return;
    static std::vector<std::vector<int>> done;
    std::vector<int> key{params.getshape(0), params.getshape(1), params.getshape(2), params.getshape(3), n,
                         default_precision()};
    if (std::find(done.begin(), done.end(), key) != done.end()) return;
    std::vector<int> shapes;
    for (int k = 0; k < n; ++k)
        for (int d = 0; d < 4; ++d) shapes.push_back(params.getshape(d));
    check(lgf_autotune_batch(context(), shapes.data(), n, default_precision(), lgf_host::autotune_steps(), nullptr,
                             nullptr),
          "lgf_autotune_batch");
    done.push_back(key);
}

// ---------------------------------------------------------------------------
// LGF/LGF.cpp:163-196
// Flows `steps` steps; with checkpoints (lgf_host::checkpoint_every) the
// recorded observables are appended to the CSVs and the configuration saved
// under its flow time every n steps before the last.
static void flow_with_checkpoints(SimulationParameters& params, FlowParameters& flowParams, GradientFlow& flowing,
                                  SU3_field& U, TopologicalCharge& topCharge, EnergyDensity& Edensity,
                                  const std::string& ensembleNum) {
    const int n = lgf_host::checkpoint_every();
    const int steps = flowParams.getflowSteps();
    // a resumed flow measures every k steps from its pickup (LGF/LGF.cpp:197-226),
    // so checkpoints on multiples of k keep the measurement times of one run
    if (n > 0 && n % flowParams.getMeasuringInterval() != 0)
        throw lgf_host::Error(LGF_E_ARG, "checkpoint interval must be a multiple of the measurement interval");
    for (int s = 1; s <= steps; s++) {
        flowing.flow();
        if (n > 0 && s % n == 0 && s < steps) {
            topCharge.saveTopologicalChargeToFile(params.getbeta(), flowing.getupdateMethod(),
                                                  flowParams.getdataFolder(), "_" + ensembleNum);
            Edensity.saveEnergyDensityToFile(params.getbeta(), flowing.getupdateMethod(), flowParams.getdataFolder(),
                                             "_" + ensembleNum);
            topCharge.clearResults();
            Edensity.clearResults();
            U.saveSU3ToFile(params.getbeta(), flowing.getupdateMethod(), flowParams.getdataFolder(),
                            ensembleNum + "_Flowtime" + std::to_string(flowing.GetFlowTime()));
        }
    }
}

void FlowSavedGaugeConfigurations(SimulationParameters& params, FlowParameters& flowParams) {
    const std::string UpdateMethod = input_method();
    Lattice lattice(params, true);
    tune_for(params, 1);
    Wilson action(params.getbeta());
    for (int i = flowParams.getConfigurationStart(); i <= flowParams.getConfigurationStop(); i++) {
        SU3_field U(lattice, params.getextdofs());
        std::string ensembleNum = std::to_string(i);
        U.loadSU3FromFile(params.getbeta(), UpdateMethod, flowParams.getdataFolder(), ensembleNum);
        GradientFlow flowing(action, U, flowParams.getepsilon(), flowParams.getMeasuringInterval());
        TopologicalCharge topCharge(U);
        flowing.Include_TopCharge(topCharge);
        topCharge.calculate(0.0);
        EnergyDensity Edensity(U);
        flowing.Include_EnergyDensity(Edensity);
        Edensity.calculate(0.0);
        flow_with_checkpoints(params, flowParams, flowing, U, topCharge, Edensity, ensembleNum);
        double flowTime = flowing.GetFlowTime();
        topCharge.saveTopologicalChargeToFile(params.getbeta(), flowing.getupdateMethod(), flowParams.getdataFolder(),
                                              "_" + ensembleNum);
        Edensity.saveEnergyDensityToFile(params.getbeta(), flowing.getupdateMethod(), flowParams.getdataFolder(),
                                         "_" + ensembleNum);
        U.saveSU3ToFile(params.getbeta(), flowing.getupdateMethod(), flowParams.getdataFolder(),
                        ensembleNum + "_Flowtime" + std::to_string(flowTime));
    }
}

// All configurations of the range flow together: one set of kernel launches
// per RK3 step for the whole batch, one clover sweep per measurement.  Each
// configuration's numbers are bitwise those of the one-at-a-time driver; the
// console lines are emitted per configuration in the reference's order.
void FlowSavedGaugeConfigurationsBatched(SimulationParameters& params, FlowParameters& flowParams) {
    const std::string UpdateMethod = input_method();
    Lattice lattice(params, true);
    const int c0 = flowParams.getConfigurationStart(), c1 = flowParams.getConfigurationStop();
    const int n = c1 - c0 + 1;
    if (n <= 0) return;
    tune_for(params, n);
    std::vector<std::unique_ptr<SU3_field>> U;
    std::vector<lgf_field*> handles;
    for (int i = c0; i <= c1; ++i) {
        U.emplace_back(new SU3_field(lattice, params.getextdofs()));
        U.back()->loadSU3FromFile(params.getbeta(), UpdateMethod, flowParams.getdataFolder(), std::to_string(i));
        handles.push_back(U.back()->handle());
    }
    lgf_batch* batch = nullptr;
    check(lgf_batch_create(context(), handles.data(), n, &batch), "lgf_batch_create");
    std::vector<TopologicalCharge> Qs;
    std::vector<EnergyDensity> Es;
    for (int k = 0; k < n; ++k) {
        Qs.emplace_back(*U[k]);
        Es.emplace_back(*U[k]);
    }
    std::vector<double> E(n), Q(n), Qa(n), eps(n, flowParams.getepsilon());
    std::vector<std::ostringstream> log(n);
    std::streambuf* console = std::cout.rdbuf();
    auto measure = [&](double t) {
        int rc = lgf_batch_measure(batch, measure_flags(), E.data(), Q.data(), Qa.data());
        if (rc != LGF_OK && rc != LGF_E_NAN) check(rc, "lgf_batch_measure");
        for (int k = 0; k < n; ++k) {
            std::cout.rdbuf(log[k].rdbuf());
            Qs[k].record(t, Q[k]);
            Es[k].record(t, E[k]);
        }
        std::cout.rdbuf(console);
    };
    double t = 0.0;
    measure(0.0);
    const int every = flowParams.getMeasuringInterval();
    for (int s = 1; s <= flowParams.getflowSteps(); ++s) {
        check(lgf_batch_flow_steps(batch, eps.data(), 1, 0), "lgf_batch_flow_steps");
        t += flowParams.getepsilon();
        if (s % every == 0) measure(t);
    }
    lgf_batch_destroy(batch);
    for (int k = 0; k < n; ++k) {
        std::cout << log[k].str();
        std::string num = std::to_string(c0 + k);
        Qs[k].saveTopologicalChargeToFile(params.getbeta(), "GF", flowParams.getdataFolder(), "_" + num);
        Es[k].saveEnergyDensityToFile(params.getbeta(), "GF", flowParams.getdataFolder(), "_" + num);
        U[k]->saveSU3ToFile(params.getbeta(), "GF", flowParams.getdataFolder(),
                            num + "_Flowtime" + std::to_string(t));
    }
}

// LGF/LGF.cpp:197-226
void ResumeFlowedConfiguration(SimulationParameters& params, FlowParameters& flowParams, double flowTime_pickup) {
    Lattice lattice(params, true);
    tune_for(params, 1);
    Wilson action(params.getbeta());
    for (int i = flowParams.getConfigurationStart(); i <= flowParams.getConfigurationStop(); i++) {
        SU3_field U(lattice, params.getextdofs());
        std::string ensembleNum = std::to_string(i);
        U.loadSU3FromFile(params.getbeta(), "GF", flowParams.getdataFolder(),
                          ensembleNum + "_Flowtime" + std::to_string(flowTime_pickup));
        GradientFlow flowing(action, U, flowParams.getepsilon(), flowParams.getMeasuringInterval());
        flowing.SetFlowTime(flowTime_pickup);
        TopologicalCharge topCharge(U);
        flowing.Include_TopCharge(topCharge);
        EnergyDensity Edensity(U);
        flowing.Include_EnergyDensity(Edensity);
        flow_with_checkpoints(params, flowParams, flowing, U, topCharge, Edensity, ensembleNum);
        topCharge.saveTopologicalChargeToFile(params.getbeta(), flowing.getupdateMethod(), flowParams.getdataFolder(),
                                              "_" + ensembleNum);
        Edensity.saveEnergyDensityToFile(params.getbeta(), flowing.getupdateMethod(), flowParams.getdataFolder(),
                                         "_" + ensembleNum);
        double NewflowTime = flowing.GetFlowTime();
        U.saveSU3ToFile(params.getbeta(), flowing.getupdateMethod(), flowParams.getdataFolder(),
                        ensembleNum + "_Flowtime" + std::to_string(NewflowTime));
    }
}

// LGF/LGF.cpp:114-124
void MeasureQdensityDistribution(SimulationParameters& params, FlowParameters& flowParams) {
    Lattice lattice(params, true);
    double flowTime_pickup = flowParams.getepsilon() * flowParams.getflowSteps();
    for (int i = flowParams.getConfigurationStart(); i <= flowParams.getConfigurationStop(); i++) {
        SU3_field U(lattice, params.getextdofs());
        std::string ensembleNum = std::to_string(i);
        U.loadSU3FromFile(params.getbeta(), "GF", flowParams.getdataFolder(),
                          ensembleNum + "_Flowtime" + std::to_string(flowTime_pickup));
        QdensityDistribution Qdensity(U, params.getbeta(), "GF", flowParams.getdataFolder(),
                                      ensembleNum + "_Flowtime" + std::to_string(flowTime_pickup));
        Qdensity.calculate(flowTime_pickup);
    }
}

// LGF/LGF.cpp:228-255: unimproved Q and improved E of the flowed
// configurations <id>_Flowtime<eps*flowSteps>, printed Q first, appended to
// the CSVs torus_extdof4_<id>.csv
void MeasureFlowedGaugeConfigurations(SimulationParameters& params, FlowParameters& flowParams) {
    Lattice lattice(params, true);
    double flowTime_pickup = flowParams.getepsilon() * flowParams.getflowSteps();
    for (int i = flowParams.getConfigurationStart(); i <= flowParams.getConfigurationStop(); i++) {
        SU3_field U(lattice, params.getextdofs());
        std::string ensembleNum = std::to_string(i);
        U.loadSU3FromFile(params.getbeta(), "GF", flowParams.getdataFolder(),
                          ensembleNum + "_Flowtime" + std::to_string(flowTime_pickup));
        UnImprovedTopologicalCharge unimprovedTopCharge(U);
        unimprovedTopCharge.calculate(flowTime_pickup);
        ImprovedEnergyDensity ImprovedEdensity(U);
        ImprovedEdensity.calculate(flowTime_pickup);
        unimprovedTopCharge.saveTopologicalChargeToFile(params.getbeta(), "GF", flowParams.getdataFolder(),
                                                        "_" + ensembleNum);
        ImprovedEdensity.saveEnergyDensityToFile(params.getbeta(), "GF", flowParams.getdataFolder(), "_" + ensembleNum);
    }
}

// LGF/LGF.cpp:308-374: same positional layout and the same rule that an
// integer in the epsilon slot means OR-per-heatbath instead.
void ParseCLargs(int argc, char** argv, int& NrDims, int& extdofs, double& beta, std::vector<int>& shape,
                 std::vector<int>& cuts, int& ConfigStart, int& ConfigStop, int& flowSteps, double& epsilon,
                 int& measureInterval, int& ThermSteps, int& OR_per_HB, int& configSep, std::string& dataFolder) {
    if (argc < 2) throw Error(LGF_E_ARG, "missing arguments");
    NrDims = std::stoi(argv[1]);
    if (argc < 10 + 2 * NrDims) throw Error(LGF_E_ARG, "expected NrDims extdofs beta shape cuts configStart "
                                                       "configStop flowSteps epsilon measureInterval dataFolder");
    extdofs = std::stoi(argv[2]);
    beta = std::stod(argv[3]);
    int i = 4;
    shape.clear();
    cuts.clear();
    for (; i < 4 + NrDims; i++) {
        shape.push_back(std::stoi(argv[i]));
        cuts.push_back(std::stoi(argv[i + NrDims]));
    }
    i += NrDims;
    ConfigStart = std::stoi(argv[i++]);
    ConfigStop = std::stoi(argv[i++]);
    flowSteps = std::stoi(argv[i]);
    ThermSteps = std::stoi(argv[i]);
    i += 1;
    std::size_t used = 0;
    std::stoi(argv[i], &used);
    if (used == std::strlen(argv[i])) {
        OR_per_HB = std::stoi(argv[i]);
        epsilon = 2147483647.0;
    } else {
        OR_per_HB = 2147483647;
        epsilon = std::stod(argv[i]);
    }
    i += 1;
    measureInterval = std::stoi(argv[i]);
    configSep = std::stoi(argv[i]);
    i += 1;
    dataFolder = argv[i];
}
