// Test-infrastructure harness around the reference LGF sources (never shipped,
// never linked into the product).  It is compiled together with the reference's
// own translation units, read in place from the reference tree, by
// oracle/ref/Makefile, and drives the reference's public entry points with the
// reference's own positional CLI parser (LGF/LGF.cpp:308-374):
//
//   lgf_ref flow      <ParseCLargs args...>   -> FlowSavedGaugeConfigurations   (LGF/LGF.cpp:163-196)
//   lgf_ref qdensity  <ParseCLargs args...>   -> MeasureQdensityDistribution    (LGF/LGF.cpp:114-124)
//   lgf_ref resume T  <ParseCLargs args...>   -> ResumeFlowedConfiguration      (LGF/LGF.cpp:197-226)
//   lgf_ref measured  <ParseCLargs args...>   -> MeasureFlowedGaugeConfigurations (LGF/LGF.cpp:228-255)
//   lgf_ref timeflow  <ParseCLargs args...>   -> the objects FlowSavedGaugeConfigurations builds for
//       configuration ConfigStart, then timed: flowSteps calls of GradientFlow::flow() with no
//       measurement due, and one TopologicalCharge + EnergyDensity calculate() (the measurement the
//       flow makes every interval); file I/O outside the timed regions.  Times are the slowest
//       rank's (barriers around each region), printed on stderr.
//
// The only behavioural change is std::cout.precision(17), so that the
// "Topological Charge = " / "Energy Density = " lines printed by the reference
// (LGF/Observables/TopologicalCharge.cpp:25, EnergyDensity.cpp:30) carry all
// significant digits.  Wall time of the whole call is printed on stderr.
#include "LGF.h"
#include "Gradient_Flow/GradientFlow.h"
#include "Observables/EnergyDensity.h"
#include "Observables/TopologicalCharge.h"
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

int main(int argc, char** argv) {
    mpiWrapper::begin_parallelSession(argc, argv);
    std::cout.precision(17);
    if (argc < 3) {
        if (mpiWrapper::id() == 0)
            std::cerr << "usage: lgf_ref flow|qdensity|measured|resume <t> <LGF positional args>\n";
        mpiWrapper::end_parallelSession();
        return 2;
    }
    std::string mode = argv[1];
    int shift = 1;
    double t_pickup = 0.0;
    if (mode == "resume") {
        t_pickup = std::stod(argv[2]);
        shift = 2;
    }
    // ParseCLargs expects argv[0] to be the program name.
    std::vector<char*> args;
    args.push_back(argv[0]);
    for (int i = 1 + shift; i < argc; ++i) args.push_back(argv[i]);

    int NrDims, extdofs;
    double beta;
    std::vector<int> shape, cuts;
    int ConfigStart, ConfigStop, flowSteps, measureInterval, ThermSteps, OR_per_HB, configSep;
    double epsilon;
    std::string dataFolder;
    ParseCLargs((int)args.size(), args.data(), NrDims, extdofs, beta, shape, cuts, ConfigStart,
                ConfigStop, flowSteps, epsilon, measureInterval, ThermSteps, OR_per_HB, configSep,
                dataFolder);
    SimulationParameters params(NrDims, extdofs, shape, beta, cuts, "torus");
    FlowParameters flowParams(ConfigStart, ConfigStop, flowSteps, epsilon, measureInterval, dataFolder);

    auto t0 = std::chrono::steady_clock::now();
    if (mode == "flow") {
        FlowSavedGaugeConfigurations(params, flowParams);
    } else if (mode == "qdensity") {
        MeasureQdensityDistribution(params, flowParams);
    } else if (mode == "measured") {
        MeasureFlowedGaugeConfigurations(params, flowParams);
    } else if (mode == "resume") {
        ResumeFlowedConfiguration(params, flowParams, t_pickup);
    } else if (mode == "timeflow") {
        Lattice lattice(params, true);
        Wilson action(params.getbeta());
        SU3_field U(lattice, params.getextdofs());
        U.loadSU3FromFile(params.getbeta(), "LHMC_3_of_0_330000", dataFolder, std::to_string(ConfigStart));
        GradientFlow flowing(action, U, epsilon, INT_MAX);
        TopologicalCharge topCharge(U);
        EnergyDensity Edensity(U);
        MPI_Barrier(mpiWrapper::comm());
        auto a = std::chrono::steady_clock::now();
        for (int i = 0; i < flowSteps; i++) flowing.flow();
        MPI_Barrier(mpiWrapper::comm());
        auto b = std::chrono::steady_clock::now();
        topCharge.calculate(flowing.GetFlowTime());
        Edensity.calculate(flowing.GetFlowTime());
        MPI_Barrier(mpiWrapper::comm());
        auto c = std::chrono::steady_clock::now();
        if (mpiWrapper::id() == 0)
            std::cerr << "lgf_ref flow_s " << std::chrono::duration<double>(b - a).count() << " steps " << flowSteps
                      << " measure_s " << std::chrono::duration<double>(c - b).count() << "\n";
    } else {
        if (mpiWrapper::id() == 0) std::cerr << "unknown mode " << mode << "\n";
        mpiWrapper::end_parallelSession();
        return 2;
    }
    auto t1 = std::chrono::steady_clock::now();
    if (mpiWrapper::id() == 0)
        std::cerr << "lgf_ref wall_s " << std::chrono::duration<double>(t1 - t0).count()
                  << " ranks " << mpiWrapper::nProcs() << "\n";
    mpiWrapper::end_parallelSession();
    return 0;
}
