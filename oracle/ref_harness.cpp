// Golden-vector harness for the Brownian/LLG step of the reference program.
//
// TEST INFRASTRUCTURE ONLY. This file is the project's own driver. It is
// compiled together with the reference's unmodified sources, read in place
// from the reference checkout (see oracle/Makefile), with the reference's
// `main` renamed away. It calls the reference's own functions in the order
// RunSimulation uses them (src/simulation.cpp:466-520) and dumps every array
// the step reads or writes, so the project's CPU restatement and HIP kernels
// can be compared against the reference on identical inputs.
//
// The reference reads ../configuration.ini at static-init time
// (src/configuration.cpp:6, src/simulation.cpp:17), so the caller runs this
// binary from a directory whose parent holds that file.
//
// Usage:
//   ref_harness <out_prefix> <alphaL> <rcFrac> <nc> <steps> [state.bin] [dump_tables]
//     alphaL, rcFrac, nc : Ewald alpha*L, r_c/L and n_c, set directly
//                          (CompEwaldParams, src/interaction.cpp:164-324, picks
//                          them by wall-clock timing and is not reproducible)
//     steps              : number of RunSimulation iterations to record
//     state.bin          : optional override of pos/posMm/posEa after init
//                          (3 column-major N x 3 double arrays back to back)
//
// Output: <out_prefix>.idx (one "name rows cols" line per array) and
//         <out_prefix>.bin (column-major little-endian doubles, concatenated);
//         the reference's own text outputs (data.txt, coords.txt, mz.txt,
//         t.txt, dtMean.txt, P*.txt, MSD*.txt) in the working directory.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

// simulation.h declares the reference's own `int main(void)`; rename it so
// this driver can define the program entry point.
#define main reference_main
#include "simulation.h"
#undef main
#include "configuration.h"

extern Rand::Vmt19937_64 generator;  // defined in src/simulation.cpp:20

namespace {

struct Dump {
    std::ofstream idx, bin;
    explicit Dump(const std::string& prefix)
        : idx(prefix + ".idx"), bin(prefix + ".bin", std::ios::binary) {}
    void put(const std::string& name, const double* p, long rows, long cols) {
        idx << name << " " << rows << " " << cols << "\n";
        bin.write(reinterpret_cast<const char*>(p), sizeof(double) * rows * cols);
    }
    void arr(const std::string& name, const ArrayXXd& a) { put(name, a.data(), a.rows(), a.cols()); }
    void vec(const std::string& name, const ArrayXd& a) { put(name, a.data(), a.rows(), 1); }
    void scal(const std::string& name, double v) { put(name, &v, 1, 1); }
};

WorkingVar_S wv;
Buffer_S buf;
IntCoefficient_S ic;
OutputVar_S ov;
Params_S prm;

}  // namespace

int main(int argc, char** argv) {
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s out_prefix alphaL rcFrac nc steps [state.bin]\n", argv[0]);
        return 2;
    }
    const std::string prefix = argv[1];
    const double alphaL = std::atof(argv[2]);
    const double rcFrac = std::atof(argv[3]);
    const int nc = std::atoi(argv[4]);
    const int steps = std::atoi(argv[5]);
    const char* statePath = argc > 6 ? argv[6] : nullptr;
    const int N = config->getNumbPart();

    // InitSimulation (src/simulation.cpp:52-105) minus the timing-based tuner.
    InitParticleProps(&wv.partProp);
    prm.lBox = std::pow(wv.partProp.volHydr.sum() / config->getVolFrac(), 1.0 / 3.0);
    prm.volBox = std::pow(prm.lBox, 3.0);
    InitCoordsFromSketch(&wv, &prm);
    InitIntCoeff(&ic);
    InitBuffers(&buf);
    InitCutOffSR(&prm, &wv.partProp);
    long maxPairs = (long)N * (N - 1) / 2;
    wv.intLists.intVecsCoords.resize(maxPairs, 3);
    wv.intLists.intVecsLen.resize(maxPairs);
    wv.intLists.intPairs.resize(2 * maxPairs);
    wv.intLists.intRangeSR.resize(maxPairs);
    // The reference leaves this list uninitialised (src/simulation.cpp:84);
    // a fresh all-false list is what a first call sees on a zeroed page.
    wv.intLists.intRangeSR.setConstant(false);
    wv.extFluxDens.resize(N, 3);
    wv.extFluxDens.setZero();
    wv.demagFluxDens.resize(N, 3);
    wv.demagFluxDens.setZero();
    wv.intForce.resize(N, 3);
    wv.intForce.setZero();
    wv.thermTorque.resize(N, 3);
    wv.thermField.resize(N, 3);
    wv.thermForce.resize(N, 3);
    ov.posRef.resize(N, 3);
    ov.truePos.resize(N, 3);

    prm.rCutLR = rcFrac * prm.lBox;
    prm.alpha = alphaL / prm.lBox;
    prm.nCutLR = nc;
    prm.transOutput = false;
    InitEwaldConst(&prm, &buf);

    if (statePath) {
        std::ifstream in(statePath, std::ios::binary);
        ArrayXXd* dst[3] = {&wv.coordsTrans.pos, &wv.coordsRot.posMm, &wv.coordsRot.posEa};
        for (ArrayXXd* a : dst) in.read(reinterpret_cast<char*>(a->data()), sizeof(double) * N * 3);
        if (!in) {
            std::fprintf(stderr, "state file too short\n");
            return 3;
        }
    }

    // Transport observables start from the state the loop starts from
    // (src/simulation.cpp:107-112); data.txt goes to the working directory.
    ov.posRef = wv.coordsTrans.pos;
    ov.truePos = wv.coordsTrans.pos;
    if (N >= 22) WriteData2TXT(&wv, &prm, "data.txt");

    Dump d(prefix);
    d.scal("N", N);
    d.scal("lBox", prm.lBox);
    d.scal("alpha", prm.alpha);
    d.scal("rCutLR", prm.rCutLR);
    d.scal("rCutSR", prm.rCutSR);
    d.scal("nc", prm.nc);
    d.vec("rMag", wv.partProp.rMag);
    d.vec("rHydr", wv.partProp.rHydr);
    d.vec("magMom", wv.partProp.magMom);
    d.vec("zetaRot", wv.partProp.zetaRot);
    d.vec("zetaTrans", wv.partProp.zetaTrans);
    d.vec("volMag", wv.partProp.volMag);
    d.arr("pos0", wv.coordsTrans.pos);
    d.arr("mm0", wv.coordsRot.posMm);
    d.arr("ea0", wv.coordsRot.posEa);
    // Closed-form Ewald error models on a grid (src/interaction.cpp:330-347).
    {
        double M = (wv.partProp.magMom.pow(2.0)).sum();
        double fAl = 3 * config->getMu0() * std::pow(wv.partProp.magMom.minCoeff(), 2.0) /
                     (2.0 * config->getMyPi() * std::pow(2.0 * wv.partProp.rHydr.minCoeff(), 4.0));
        ArrayXXd g(40, 3);
        for (int i = 0; i < 40; ++i) {
            double x = (i + 1) * 2.5e6;      // alpha grid (1/m)
            double kc = 0.5 + 0.5 * i;       // k_c grid
            g(i, 0) = x;
            g(i, 1) = FunRealError(x, M, fAl, &prm);
            g(i, 2) = FunRezError(kc, M, fAl, &prm);
        }
        d.arr("errgrid", g);
        d.scal("errM", M);
        d.scal("errFAl", fAl);
    }

    double tInt = 0.0;
    double deltaT = config->getDeltaTInit();
    double deltaTused = 0.0;
    int countCoords = 0;
    for (int s = 0; s < steps; ++s) {
        const std::string p = "s" + std::to_string(s) + "_";
        d.arr(p + "pos_in", wv.coordsTrans.pos);
        d.arr(p + "mm_in", wv.coordsRot.posMm);
        d.arr(p + "ea_in", wv.coordsRot.posEa);
        d.scal(p + "t_in", tInt);
        d.scal(p + "dt_in", deltaT);

        DefExtField(&wv.extFluxDens, tInt);
        DefThermFluct(&wv, &buf.buffer1d, deltaT);
        d.arr(p + "nT", wv.thermTorque);
        d.arr(p + "nB", wv.thermField);
        d.arr(p + "nF", wv.thermForce);

        EvalSinCos(&wv, &buf, &prm);
        ov.Pxy = ov.Pxz = ov.Pyz = 0;
        CompInteractions(&wv, &buf, &prm, &ov);
        d.arr(p + "force", wv.intForce);
        d.arr(p + "field", wv.demagFluxDens);
        d.arr(p + "pos_mid", wv.coordsTrans.pos);
        {
            double P[3] = {ov.Pxy, ov.Pxz, ov.Pyz};
            d.put(p + "press", P, 3, 1);
        }
        d.scal(p + "pairs", wv.intLists.pairs);

        IntegrationRot(&wv, &ic, &buf, &prm, &deltaT, &deltaTused);
        d.arr(p + "mm_out", wv.coordsRot.posMm);
        d.arr(p + "ea_out", wv.coordsRot.posEa);
        d.arr(p + "field_out", wv.demagFluxDens);
        d.arr(p + "nT_out", wv.thermTorque);
        d.arr(p + "nB_out", wv.thermField);
        d.scal(p + "dt_out", deltaT);
        d.scal(p + "dt_used", deltaTused);

        IntegrationTrans(&wv, &buf, &prm, deltaTused);
        ApplyBoundaryCondArr(&wv.coordsTrans.pos, &buf.buffer3d, &prm);
        d.arr(p + "pos_out", wv.coordsTrans.pos);
        d.scal(p + "mz", wv.coordsRot.posMm.col(2).mean());

        // RunSimulation's post-processing (src/simulation.cpp:490-516)
        if (std::floor(tInt / config->getTEnd() * config->getNPCoords()) == countCoords) {
            countCoords++;
            WriteCoords2TXT(&wv, "coords.txt", tInt);
        }
        EvalTransCoeffs(&wv, &buf, &ov, &prm);
        ov.t.push_back(tInt);
        ov.mz.push_back(wv.coordsRot.posMm.col(2).mean());
        {
            double M[3] = {ov.MSDx.back(), ov.MSDy.back(), ov.MSDz.back()};
            d.put(p + "msd", M, 3, 1);
            double P[3] = {ov.Pxy, ov.Pxz, ov.Pyz};
            d.put(p + "press_vol", P, 3, 1);
        }
        d.arr(p + "true_pos", ov.truePos);
        tInt += deltaTused;
    }
    d.scal("t_end", tInt);
    // end-of-run exports (src/simulation.cpp:528-537)
    if (steps > 0) {
        WriteCoords2TXT(&wv, "coords.txt", tInt);
        ExportDiffVis(&ov, tInt, steps);
        saveData("mz.txt", Map<VectorXd, Unaligned>(ov.mz.data(), ov.mz.size()), 0);
        saveData("t.txt", Map<VectorXd, Unaligned>(ov.t.data(), ov.t.size()), 0);
    }
    return 0;
}
