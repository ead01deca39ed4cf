// Command-line driver, positional arguments as the reference's
// (LGF/LGF.cpp:20-37, 308-374):
//   lgf NrDims extdofs beta N0 N1 N2 N3 c0 c1 c2 c3 configStart configStop
//       flowSteps epsilon measureInterval dataFolder  [options]
// The reference picks its entry point by editing main(); here an option does:
//   --mode=flow        FlowSavedGaugeConfigurations (default)
//   --mode=batch       the same, all configurations of the range in one batch
//   --mode=qdensity    MeasureQdensityDistribution
//   --mode=resume --pickup=T   ResumeFlowedConfiguration
//   --mode=measured    MeasureFlowedGaugeConfigurations
//   --checkpoint-every=N   (flow, resume) every N steps, a multiple of the
//                      measurement interval, append the observables so far to
//                      the CSVs and save the configuration as <id>_Flowtime<t>
//   --autotune[=N]     before flowing, time every tile order and stage variant
//                      on the run's lattice shape for N RK3 steps (default 5)
//                      and keep the fastest (launch tuning only; FP64 results
//                      are unchanged)
//   --precision=fp64|fp32      flow arithmetic (fp64 default)
//   --measure=fp64|native      clover arithmetic of measurements (fp64 default)
//   --input-method=NAME        input update-method directory (default LHMC_3_of_0_330000)
//   --device=N
// Errors print a message and exit non-zero; nothing waits on stdin.
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "lgf_host.hpp"

int main(int argc, char** argv) {
    std::vector<char*> pos;
    pos.push_back(argv[0]);
    std::string mode = "flow";
    double pickup = 0.0;
    try {
        for (int i = 1; i < argc; ++i) {
            std::string a = argv[i];
            if (a.rfind("--", 0) != 0) {
                pos.push_back(argv[i]);
                continue;
            }
            auto val = [&](const char* key) { return a.substr(std::strlen(key)); };
            if (a.rfind("--mode=", 0) == 0) mode = val("--mode=");
            else if (a.rfind("--pickup=", 0) == 0) pickup = std::stod(val("--pickup="));
            else if (a == "--precision=fp64") lgf_host::set_default_precision(LGF_FP64);
            else if (a == "--precision=fp32") lgf_host::set_default_precision(LGF_FP32);
            else if (a == "--measure=fp64") lgf_host::set_measure_flags(LGF_MEASURE_FP64);
            else if (a == "--measure=native") lgf_host::set_measure_flags(LGF_MEASURE_DEFAULT);
            else if (a.rfind("--input-method=", 0) == 0) lgf_host::set_input_method(val("--input-method="));
            else if (a.rfind("--device=", 0) == 0) lgf_host::set_device(std::stoi(val("--device=")));
            else if (a == "--autotune") lgf_host::set_autotune(5);
            else if (a.rfind("--autotune=", 0) == 0) lgf_host::set_autotune(std::stoi(val("--autotune=")));
            else if (a.rfind("--checkpoint-every=", 0) == 0)
                lgf_host::set_checkpoint_every(std::stoi(val("--checkpoint-every=")));
            else throw lgf_host::Error(LGF_E_ARG, "unknown option " + a);
        }
        int NrDims, extdofs, ConfigStart, ConfigStop, flowSteps, measureInterval, ThermSteps, OR_per_HB, configSep;
        double beta, epsilon;
        std::vector<int> shape, cuts;
        std::string dataFolder;
        ParseCLargs((int)pos.size(), pos.data(), NrDims, extdofs, beta, shape, cuts, ConfigStart, ConfigStop,
                    flowSteps, epsilon, measureInterval, ThermSteps, OR_per_HB, configSep, dataFolder);
        SimulationParameters params(NrDims, extdofs, shape, beta, cuts, "torus");
        FlowParameters flowParams(ConfigStart, ConfigStop, flowSteps, epsilon, measureInterval, dataFolder);
        if (mode == "flow") FlowSavedGaugeConfigurations(params, flowParams);
        else if (mode == "batch") FlowSavedGaugeConfigurationsBatched(params, flowParams);
        else if (mode == "qdensity") MeasureQdensityDistribution(params, flowParams);
        else if (mode == "resume") ResumeFlowedConfiguration(params, flowParams, pickup);
        else if (mode == "measured") MeasureFlowedGaugeConfigurations(params, flowParams);
        else throw lgf_host::Error(LGF_E_ARG, "unknown mode " + mode);
    } catch (const lgf_host::Error& e) {
        std::cerr << "lgf: " << e.what() << "\n";
        return 10 + e.code;
    } catch (const std::exception& e) {
        std::cerr << "lgf: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
