// C++ host API of the flow path, mirroring the reference LGF driver so that
// code written against it keeps compiling and behaving the same, while the
// work runs on the GPU through the C-ABI of include/lgf_flow.h.
//
// Kept from the reference (same names, constructor arguments and meaning):
//   SimulationParameters, FlowParameters     LGF/SimulationParameters.h:3-86
//   Lattice                                  LGF/Lattice/Lattice.h (shape/volume/type only)
//   SU3_field (the gauge field)              LGF/Field/SU3_field.h:7-39 (load/save, cold start)
//   Wilson                                   LGF/Action/Gauge/Wilson.h:7
//   GradientFlow                             LGF/Gradient_Flow/GradientFlow.h:7-34
//   EnergyDensity, TopologicalCharge         LGF/Observables/{EnergyDensity,TopologicalCharge}.h
//   ImprovedEnergyDensity,                   LGF/Observables/ImprovedEnergyDensity.h,
//   UnImprovedTopologicalCharge              LGF/Observables/UnImprovedTopologicalCharge.h
//   QdensityDistribution                     LGF/Observables/QdensityDistribution.h
//   FlowSavedGaugeConfigurations, MeasureQdensityDistribution,
//   ResumeFlowedConfiguration, MeasureFlowedGaugeConfigurations,
//   ParseCLargs                              LGF/LGF.h:29-40
// Differences, all deliberate:
//   * errors throw lgf_host::Error instead of exit()/blocking on std::cin
//     (LGF/SimulationParameters.h:12-19, LGF/Field/SU3_field.cpp:299-309);
//   * "torus_OpenT" is rejected (periodic lattices only); `cuts` is accepted
//     and ignored (one GPU holds the whole lattice);
//   * configuration offsets are 64-bit (LGF/Field/SU3_field.cpp:331 uses int);
//   * every CSV also gets a full-precision twin "<name>.full.csv".
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lgf_flow.h"

namespace lgf_host {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// process-wide device context used by all host objects (device 0 unless
// set_device() is called before the first object is created)
lgf_ctx* context();
void set_device(int device);
void check(int rc, const char* what);

// precision used for the flow of new SU3_field objects (LGF_FP64 default)
void set_default_precision(int precision);
int default_precision();
// clover arithmetic of measurements (LGF_MEASURE_FP64 default)
void set_measure_flags(int flags);
int measure_flags();
// Launch tuning of the flow (0 = off, the default): before the first flow of
// a lattice shape, the drivers time every tile order and stage variant for
// `nsteps` RK3 steps on seeded lattices of that shape in the flow precision
// (lgf_autotune_batch; as many lattices as the batched driver flows at
// once) and keep the fastest.  FP64 results do not depend on the choice;
// FP32 results can differ from the untuned ones by FP32 rounding.
void set_autotune(int nsteps);
int autotune_steps();

}  // namespace lgf_host

class SimulationParameters {
public:
    SimulationParameters(int NrDims, int extdofs, std::vector<int> shape, double beta, std::vector<int> cuts,
                         std::string LatticeType = "torus");
    int getNrDims() { return m_NrDims; }
    int getextdofs() { return m_extdofs; }
    int getshape(int i) { return m_shape[i]; }
    double getbeta() { return m_beta; }
    int getcuts(int i) { return m_cuts[i]; }
    std::string getLatticeType() { return m_LatticeType; }

private:
    int m_NrDims;
    int m_extdofs;
    std::vector<int> m_shape;
    double m_beta;
    std::vector<int> m_cuts;
    std::string m_LatticeType;
};

class FlowParameters {
public:
    FlowParameters(int ConfigurationStart, int ConfigurationStop, int flowSteps, double epsilon,
                   int measure_every_nth_step, const std::string& dataFolder)
        : m_ConfigurationStart(ConfigurationStart),
          m_ConfigurationStop(ConfigurationStop),
          m_flowSteps(flowSteps),
          m_epsilon(epsilon),
          m_measure_every_nth_step(measure_every_nth_step),
          m_dataFolder(dataFolder) {}
    int getConfigurationStart() { return m_ConfigurationStart; }
    int getConfigurationStop() { return m_ConfigurationStop; }
    int getflowSteps() { return m_flowSteps; }
    double getepsilon() { return m_epsilon; }
    int getMeasuringInterval() { return m_measure_every_nth_step; }
    std::string getdataFolder() { return m_dataFolder; }

private:
    int m_ConfigurationStart;
    int m_ConfigurationStop;
    int m_flowSteps;
    double m_epsilon;
    int m_measure_every_nth_step;
    std::string m_dataFolder;
};

class Lattice {
public:
    Lattice(SimulationParameters& params, bool cloversRequired);
    int getNdims() const { return 4; }
    const int* getShape() const { return m_shape; }
    std::string getType() const { return m_type; }
    long long m_totalVolume;

private:
    int m_shape[4];
    std::string m_type;
};

class SU3_field {
public:
    SU3_field(Lattice& lattice, int NrExtDOF);
    ~SU3_field();
    SU3_field(const SU3_field&) = delete;
    SU3_field& operator=(const SU3_field&) = delete;
    void loadSU3FromFile(double beta, const std::string& updateMethod, const std::string& dataFolder,
                         const std::string& identifier = "");
    void saveSU3ToFile(double beta, const std::string& updateMethod, const std::string& dataFolder,
                       const std::string& identifier = "");
    void InitializeColdStart();
    double Avg_Plaquette();
    Lattice& getLatticePtr() const { return *m_lattice; }
    int getNrExtDOF() const { return 4; }
    lgf_field* handle() const { return m_field; }
    int precision() const { return m_precision; }

private:
    Lattice* m_lattice;
    lgf_field* m_field = nullptr;
    int m_precision;
};

class Action {
public:
    virtual ~Action() = default;
};

class Wilson : public Action {
public:
    explicit Wilson(double beta) : m_beta(beta) {}
    double getBeta() const { return m_beta; }

private:
    double m_beta;
};

class TopologicalCharge;
class EnergyDensity;

class GradientFlow {
public:
    GradientFlow(Action& GaugeAction, SU3_field& U, double epsilon, int FlowStepsPerMeasurement = 1);
    void flow();
    void Include_TopCharge(TopologicalCharge& TopCharge) { m_topologicalCharge = &TopCharge; }
    void Include_EnergyDensity(EnergyDensity& Edensity) { m_EnergyDensity = &Edensity; }
    void MakeMeasurements();
    std::string getupdateMethod() const { return "GF"; }
    double GetFlowTime() const { return m_flowTime; }
    void SetFlowTime(double flowTime) { m_flowTime = flowTime; }

private:
    SU3_field* m_U;
    int m_NrSteps = 0;
    int m_FlowStepsPerMeasurement;
    double m_epsilon;
    double m_flowTime = 0.0;
    TopologicalCharge* m_topologicalCharge = nullptr;
    EnergyDensity* m_EnergyDensity = nullptr;
};

// One clover sweep gives E and Q together; both observables of a field share
// it, so the reference's Q-then-E print order costs one measurement.
class FieldMeasurement {
public:
    static void measure(SU3_field& U, double flowTime, double* E, double* Q, double* Qabs);
    // all lgf_observable values of the field's current links (obs[LGF_NUM_OBS])
    static void observables(SU3_field& U, double* obs);
};

class EnergyDensity {
public:
    explicit EnergyDensity(SU3_field& U) : m_U(&U) {}
    void calculate(double flowTime);
    void record(double flowTime, double value);
    void saveEnergyDensityToFile(double beta, const std::string& updateMethod, const std::string& dataFolder,
                                 const std::string& identifier = "");
    // drop the recorded series (after a checkpoint appended it to the file)
    void clearResults() {
        m_resultVector.clear();
        m_FlowMeasurementTimeVector.clear();
    }
    const std::vector<double>& values() const { return m_resultVector; }
    const std::vector<double>& times() const { return m_FlowMeasurementTimeVector; }

private:
    SU3_field* m_U;
    std::vector<double> m_resultVector;
    std::vector<double> m_FlowMeasurementTimeVector;
};

class TopologicalCharge {
public:
    explicit TopologicalCharge(SU3_field& U) : m_U(&U) {}
    void calculate(double flowTime);
    void record(double flowTime, double value);
    void saveTopologicalChargeToFile(double beta, const std::string& updateMethod, const std::string& dataFolder,
                                     const std::string& identifier = "");
    // drop the recorded series (after a checkpoint appended it to the file)
    void clearResults() {
        m_resultVector.clear();
        m_FlowMeasurementTimeVector.clear();
    }
    const std::vector<double>& values() const { return m_resultVector; }
    const std::vector<double>& times() const { return m_FlowMeasurementTimeVector; }

private:
    SU3_field* m_U;
    std::vector<double> m_resultVector;
    std::vector<double> m_FlowMeasurementTimeVector;
};

class ImprovedEnergyDensity {
public:
    explicit ImprovedEnergyDensity(SU3_field& U) : m_U(&U) {}
    void calculate(double flowTime);
    void saveEnergyDensityToFile(double beta, const std::string& updateMethod, const std::string& dataFolder,
                                 const std::string& identifier = "");
    const std::vector<double>& values() const { return m_resultVector; }

private:
    SU3_field* m_U;
    std::vector<double> m_resultVector;
    std::vector<double> m_FlowMeasurementTimeVector;
};

class UnImprovedTopologicalCharge {
public:
    explicit UnImprovedTopologicalCharge(SU3_field& U) : m_U(&U) {}
    void calculate(double flowTime);
    void saveTopologicalChargeToFile(double beta, const std::string& updateMethod, const std::string& dataFolder,
                                     const std::string& identifier = "");
    const std::vector<double>& values() const { return m_resultVector; }

private:
    SU3_field* m_U;
    std::vector<double> m_resultVector;
    std::vector<double> m_FlowMeasurementTimeVector;
};

class QdensityDistribution {
public:
    QdensityDistribution(SU3_field& U, double beta, const std::string& updateMethod, const std::string& dataFolder,
                         const std::string& identifier = "");
    void calculate(double flowTime);

private:
    SU3_field* m_U;
    double m_beta;
    std::string m_updateMethod, m_dataFolder, m_identifier;
};

void FlowSavedGaugeConfigurations(SimulationParameters& params, FlowParameters& flowParams);
void MeasureQdensityDistribution(SimulationParameters& params, FlowParameters& flowParams);
void ResumeFlowedConfiguration(SimulationParameters& params, FlowParameters& flowParams, double flowTime_pickup);
void MeasureFlowedGaugeConfigurations(SimulationParameters& params, FlowParameters& flowParams);
// FlowSavedGaugeConfigurations with all configurations of the range flowed
// together in one batch (same results per configuration, same files).
void FlowSavedGaugeConfigurationsBatched(SimulationParameters& params, FlowParameters& flowParams);
void ParseCLargs(int argc, char** argv, int& NrDims, int& extdofs, double& beta, std::vector<int>& shape,
                 std::vector<int>& cuts, int& ConfigStart, int& ConfigStop, int& flowSteps, double& epsilon,
                 int& measureInterval, int& ThermSteps, int& OR_per_HB, int& configSep, std::string& dataFolder);

namespace lgf_host {
// input update-method directory of FlowSavedGaugeConfigurations
// (the reference hard-codes "LHMC_3_of_0_330000", LGF/LGF.cpp:164)
void set_input_method(const std::string& m);
std::string input_method();
// Mid-flow checkpoints of FlowSavedGaugeConfigurations and
// ResumeFlowedConfiguration (0 = none): every n steps the observables
// recorded so far are appended to their CSVs and the configuration is saved
// as <id>_Flowtime<t>, the file ResumeFlowedConfiguration(..., t) reads.  A
// run stopped after a checkpoint resumes from it and ends with the same
// files as an uninterrupted run (the CSVs are append-only).
void set_checkpoint_every(int n);
int checkpoint_every();
// <folder>ensembles/su3/beta<b>/<N0>X..X<N3>/<method>/torus_extdof4_<id>.bin
std::string config_path(double beta, const int* shape, const std::string& method, const std::string& folder,
                        const std::string& id);
std::string beta_string(double beta);
}  // namespace lgf_host
