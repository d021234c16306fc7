// ORACLE — test infrastructure only (see voronoi_oracle.cpp).
//
// FIRE minimization, GPU branch of the reference EnergyMinimizerFIRE
// (EnergyMinimizerFIRE2D.cpp:24-41 defaults, :93-114 velocityVerletGPU,
// :156-207 fireStepGPU, :269-284 minimize), over any model with N degrees of
// freedom. The model supplies
//   int dof();                          number of degrees of freedom
//   void forces(std::vector<D2> &F);    geometry + forces at the current positions
//   void move(const std::vector<D2> &d);  x += d, wrapped into the box
//   bool topology();                    enforceTopology; false on failure
// Sums run sequentially in degree-of-freedom order.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

namespace fireo {

// one iteration's FIRE decision, for pinning the update rules
struct IterLog {
    double power, dt_before, dt_after, alpha_before, alpha_after;
    int nsince_before, nsince_after, zeroed;
    double vv_after;  // sum |v|^2 after the step's velocity update
};

template <class D2> struct State {
    std::vector<IterLog> *log = nullptr;  // when set, one entry per iteration
    int iterations = 0, maxIterations = 1000, NMin = 5, NSince = 0;
    double forceMax = 100., forceCutoff = 1e-7, alphaStart = 0.99, alpha = 0.99, dt = 0.01, dtMin = 1e-4,
           dtMax = 0.1, dtInc = 1.05, dtDec = 0.95, alphaDec = 0.9, Power = 0;
    std::vector<D2> v, f;

    void set(int maxIt, double fc, double a0, double dt0, double dtMax_, double dtInc_, double dtDec_,
             double alphaDec_, int nMin) {
        maxIterations = maxIt;
        forceCutoff = fc;
        alphaStart = a0;
        alpha = a0;
        dt = dt0;
        dtMin = dt0 * .01;
        dtMax = dtMax_;
        dtInc = dtInc_;
        dtDec = dtDec_;
        alphaDec = alphaDec_;
        NMin = nMin;
    }
};

// Runs the minimize loop; returns the iterations done so far (they accumulate
// over calls, as in the reference), or -1 when the topology step fails.
// out: forceMax (= sum |F|^2 / N), Power, alpha, dt, max |F|^2.
template <class D2, class M> int minimize(State<D2> &fr, M &model, double *out) {
    const int N = model.dof();
    if ((int)fr.v.size() != N) fr.v.assign(N, D2{0, 0});
    model.forces(fr.f);
    fr.forceMax = 110.0;
    double fmax2 = 0;
    std::vector<D2> disp(N);
    while (fr.iterations < fr.maxIterations && std::sqrt(fr.forceMax) > fr.forceCutoff) {
        fr.iterations += 1;
        // velocityVerletGPU
        for (int i = 0; i < N; ++i) {
            disp[i].x = fr.dt * fr.v[i].x + 0.5 * fr.dt * fr.dt * fr.f[i].x;
            disp[i].y = fr.dt * fr.v[i].y + 0.5 * fr.dt * fr.dt * fr.f[i].y;
        }
        for (int i = 0; i < N; ++i) {
            fr.v[i].x += 0.5 * fr.dt * fr.f[i].x;
            fr.v[i].y += 0.5 * fr.dt * fr.f[i].y;
        }
        model.move(disp);
        if (!model.topology()) return -1;
        model.forces(fr.f);
        for (int i = 0; i < N; ++i) {
            fr.v[i].x += 0.5 * fr.dt * fr.f[i].x;
            fr.v[i].y += 0.5 * fr.dt * fr.f[i].y;
        }
        // fireStepGPU
        double ff = 0, fv = 0, vv = 0;
        fmax2 = 0;
        for (int i = 0; i < N; ++i) {
            double a = fr.f[i].x * fr.f[i].x + fr.f[i].y * fr.f[i].y;
            ff += a;
            fmax2 = std::max(fmax2, a);
            fv += fr.f[i].x * fr.v[i].x + fr.f[i].y * fr.v[i].y;
            vv += fr.v[i].x * fr.v[i].x + fr.v[i].y * fr.v[i].y;
        }
        fr.Power = fv;
        fr.forceMax = ff / (double)N;
        double scaling = 0.0;
        if (ff > 0.) scaling = std::sqrt(vv / ff);
        for (int i = 0; i < N; ++i) {
            fr.v[i].x = (1 - fr.alpha) * fr.v[i].x + fr.alpha * scaling * fr.f[i].x;
            fr.v[i].y = (1 - fr.alpha) * fr.v[i].y + fr.alpha * scaling * fr.f[i].y;
        }
        IterLog lg{fr.Power, fr.dt, 0, fr.alpha, 0, fr.NSince, 0, 0, 0};
        if (fr.Power > 0) {
            if (fr.NSince > fr.NMin) {
                fr.dt = std::min(fr.dt * fr.dtInc, fr.dtMax);
                fr.alpha = fr.alpha * fr.alphaDec;
            }
            fr.NSince += 1;
        } else {
            fr.dt = fr.dt * fr.dtDec;
            fr.alpha = fr.alphaStart;
            for (int i = 0; i < N; ++i) fr.v[i] = D2{0, 0};
            lg.zeroed = 1;
        }
        if (fr.log) {
            lg.dt_after = fr.dt;
            lg.alpha_after = fr.alpha;
            lg.nsince_after = fr.NSince;
            for (int i = 0; i < N; ++i) lg.vv_after += fr.v[i].x * fr.v[i].x + fr.v[i].y * fr.v[i].y;
            fr.log->push_back(lg);
        }
    }
    if (out) {
        out[0] = fr.forceMax;
        out[1] = fr.Power;
        out[2] = fr.alpha;
        out[3] = fr.dt;
        out[4] = fmax2;
    }
    return fr.iterations;
}

}  // namespace fireo
