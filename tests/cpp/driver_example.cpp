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

    // MetaD-HMC (SU3_main.cpp:469-546) and two-stream tempering (SU3_main.cpp:549-621)
    GaugeField U_copy(gpu, 8, 8, 8, 8);
    MatrixField P(gpu, 8, 8, 8, 8), F(gpu, 8, 8, 8, 8);
    MetaBiasPotential Bias(-8.0, 8.0, 800, 0.05, 0.08, 10.0, 1000.0, true);
    CollectiveVariable CV(gpu, 8, 8, 8, 8, 4, 0.12);
    HMCMetaDKernel HMC_MetaD(U, U_copy, P, F, Bias, CV, LT_INT_OMF2, 6.0, 1.0);
    const bool accepted = HMC_MetaD(10);
    ReplicaBatch Temper(gpu, {{8, 8, 8, 8}, {8, 8, 8, 8}}, 4, 0.12);
    Temper.SetLadder({0.0, 1.0});
    Temper.SetBias(Bias);
    Temper.UseGraph(true);
    Temper.Step(10, 6.0, LT_INT_OMF2, 1.0, true, LT_SWAP_EVEN);
    std::printf("MetaD accepted %d dH %.6f labels %d %d\n", (int)accepted, HMC_MetaD.delta_H, Temper.Labels()[0],
                Temper.Labels()[1]);

    // the two-stream loop body of SU3_main.cpp:600-616 with the reference's own kernels
    GaugeField U_temper(gpu, 8, 8, 8, 8);
    U_temper.SetToHaar(3);
    HMCMetaDKernel HMC_temper(U_temper, U_copy, P, F, Bias, CV, LT_INT_OMF2_OMF4, 6.0, 1.0, true, "hmc_log.txt");
    MetadynamicsTemperingKernel ParallelTemperingSwap(U, U_temper, Bias, CV, 7);
    HeatbathKernel HeatbathLW{U, 6.0, LuescherWeiszAction(6.0)};
    OverrelaxationSubgroupKernel OverrelaxationLW{U, LuescherWeiszAction(6.0)};
    for (int n_count = 0; n_count < 2; ++n_count) {
        Iterator::Checkerboard4(HeatbathLW, 1);
        Iterator::Checkerboard4(OverrelaxationLW, 4);
        HMC_temper(4, true);
        std::printf("swap %d\n", (int)ParallelTemperingSwap());
    }
    ParityUpdate(U, U_copy);
    MatrixField Chain(gpu, 8, 8, 8, 8);
    const EnergyDensities E = EnergyDensity(U);
    std::printf("S_sym %.9f Q_plaq %.9f W2 %.9f W4 %.9f P %.9f E %.9f\n",
                ActionNormalized(U, LuescherWeiszAction(6.0)), TopChargePlaquette(U), WilsonLoop(U, Chain, 0, 2, true),
                WilsonLoop(U, Chain, 2, 4, false), std::abs(PolyakovLoop(U)), E.clover);
    CloverDerivative(U, F, 1.0);
    std::printf("%zu schedule events\n", IntegratorSchedule(LT_INT_LEAPFROG_OMF4, 1.0, 4).size());

    // rotating checkpoints and the bias file (SU3_main.cpp:287-294, 503-504, 631)
    CheckpointManager Checkpointer("checkpoints", 3, true);
    PhiloxState prng{U};
    Checkpointer.AlternatingCheckpoints(SaveConfigBMWFull, prng, U, "config.conf", "prng_state.txt",
                                        "distribution_state.txt");
    Bias.SaveParameters("metapotential.txt", true);
    Bias.SavePotential("metapotential.txt");
    const bool resumed = Checkpointer.LoadCheckpoint(LoadConfigBMWFull, prng, U, "config.conf", "prng_state.txt",
                                                     "distribution_state.txt");
    std::printf("resumed %d\n", (int)resumed);
    return 0;
}
