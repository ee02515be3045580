// Reference-style driver over the C++ layer (include/lettuce_hip.hpp): the update
// loop of SU3_main.cpp:342-466 with whole GPU sweeps.  Built by tests/test_abi.py.
#include <cstdio>

#include "lettuce_hip.hpp"

int main()
{
    using namespace lettuce_hip;
    Context gpu(0);
    GaugeField U(gpu, 8, 8, 8, 8);
    U.SetToHaar(0);
    U.SeedPRNG(1);
    HeatbathKernel Heatbath{U, 6.0};
    OverrelaxationSubgroupKernel Overrelaxation{U};
    for (int n = 0; n < 5; ++n) {
        Iterator::Checkerboard(Heatbath, 1);
        Iterator::Checkerboard(Overrelaxation, 4);
    }
    GaugeField W(gpu, 8, 8, 8, 8);
    GradientFlowKernel Flow(U, W, LT_FLOW_RK3, 0.01);
    Flow(10);
    std::printf("plaquette %.12f Q %.12f\n", PlaquetteSum(U) / (18.0 * U.Volume()), TopChargeClover(W));
    return 0;
}
