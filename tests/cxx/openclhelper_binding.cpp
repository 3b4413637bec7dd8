// The langmuirCore side of the C-ABI as a maintainer would write it: an
// OpenClHelper-shaped class over liblangmuir_gpu.so that stages the charges and
// targets exactly as OpenClHelper::launchCoulombKernel2 does
// (src/langmuirCore/openclhelper.cpp:377-424) -- [electrons, holes, charged
// defects] as charges, then the electrons' and holes' future sites -- and serves
// getOutputHost / getOutputHostFuture (:567-583); ChargeAgent::coulombGPU
// (chargeagent.cpp:273-296) assembles dE from them.
//
// Input (text): X Y Z cutoff hopping_range defects_charge prefactor inverse_kt
// sigma binding, then nE nH nD, then the electron sites, hole sites, defect
// sites, electron future sites, hole future sites.  Output: one line per charge
// entry "o[id] o[id + offset]" for the carriers, then one dE per carrier.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "langmuir_gpu.h"

class GpuHelper {   // the part of OpenClHelper that Simulation and ChargeAgent call
public:
    explicit GpuHelper(const lmg_params& p) : m_p(p) {
        okCL = lmg_create(&m_p, &m_ctx) == LMG_OK;   // openclhelper.cpp:176-185: no device, no GPU path
    }
    ~GpuHelper() { lmg_destroy(m_ctx); }
    bool okCL = false;

    // launchCoulombKernel2 / launchGaussKernel2 with the World's lists passed in
    bool launchCoulombKernel2(const std::vector<int>& e, const std::vector<int>& h, const std::vector<int>& d,
                              const std::vector<int>& ef, const std::vector<int>& hf) {
        m_s.clear();
        m_q.clear();
        for (int s : e) { m_s.push_back(s); m_q.push_back(-1); }   // setOpenCLID(i)
        for (int s : h) { m_s.push_back(s); m_q.push_back(+1); }   // setOpenCLID(nE + i)
        if (m_p.defects_charge != 0)
            for (int s : d) { m_s.push_back(s); m_q.push_back(m_p.defects_charge); }
        m_offset = int(m_s.size());
        for (int s : ef) { m_s.push_back(s); m_q.push_back(-1); }
        for (int s : hf) { m_s.push_back(s); m_q.push_back(+1); }
        m_o.assign(m_s.size(), 0.0);
        const int flags = m_p.gaussian_sigma > 0 ? LMG_GAUSS : 0;
        const lmg_status st = lmg_pair_potential(m_ctx, m_s.data(), m_q.data(), m_offset, m_s.data(),
                                                 int32_t(m_s.size()), flags, m_o.data());
        if (st != LMG_OK) std::fprintf(stderr, "lmg_pair_potential: %s\n", lmg_last_error(m_ctx));
        return st == LMG_OK;
    }
    double getOutputHost(int id) const { return m_o[id]; }
    double getOutputHostFuture(int id) const { return m_o[id + m_offset]; }

private:
    lmg_params m_p;
    lmg_ctx* m_ctx = nullptr;
    std::vector<int32_t> m_s, m_q;
    std::vector<double> m_o;
    int m_offset = 0;
};

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    FILE* f = std::fopen(argv[1], "r");
    if (!f) return 2;
    lmg_params p = {};
    int ne = 0, nh = 0, nd = 0;
    if (std::fscanf(f, "%d %d %d %d %d %d %lf %lf %lf %lf %d %d %d", &p.grid_x, &p.grid_y, &p.grid_z, &p.cutoff,
                    &p.hopping_range, &p.defects_charge, &p.prefactor, &p.inverse_kt, &p.gaussian_sigma,
                    &p.exciton_binding, &ne, &nh, &nd) != 13)
        return 2;
    auto read = [&](int n) {
        std::vector<int> v(n);
        for (int& x : v)
            if (std::fscanf(f, "%d", &x) != 1) std::exit(2);
        return v;
    };
    const std::vector<int> e = read(ne), h = read(nh), d = read(nd), ef = read(ne), hf = read(nh);
    std::fclose(f);

    GpuHelper ocl(p);
    if (!ocl.okCL) {
        std::fprintf(stderr, "no device: %s\n", lmg_last_error(nullptr));
        return 3;
    }
    if (!ocl.launchCoulombKernel2(e, h, d, ef, hf)) return 4;
    const int V = p.grid_x * p.grid_y * p.grid_z;
    std::vector<char> eg(V + 64, 0), hg(V + 64, 0);   // occupancy for bindingPotential
    for (int s : e) eg[s] = 1;
    for (int s : h) hg[s] = 1;
    // sI[1][0][0] = prefactor * iR[1][0][0] * eR[1][0][0] (potential.cpp:316-324)
    const double sI = p.prefactor * (p.gaussian_sigma > 0 ? std::erf(1.0 / (std::sqrt(2.0) * p.gaussian_sigma)) : 1.0);
    auto binding = [&](int q, int site) {   // chargeagent.cpp:463-481
        if (site < 0 || site >= V) return 0.0;
        if (q > 0) return eg[site] ? sI - p.exciton_binding : 0.0;
        return hg[site] ? sI + p.exciton_binding : 0.0;
    };
    for (int id = 0; id < ne + nh; ++id) std::printf("%.17g %.17g\n", ocl.getOutputHost(id), ocl.getOutputHostFuture(id));
    for (int id = 0; id < ne + nh; ++id) {   // ChargeAgent::coulombGPU
        const bool isE = id < ne;
        const int q = isE ? -1 : 1;
        const int site = isE ? e[id] : h[id - ne], fsite = isE ? ef[id] : hf[id - ne];
        double p1 = ocl.getOutputHost(id), p2 = ocl.getOutputHostFuture(id);
        p2 -= sI * q;
        p2 += binding(q, fsite);
        p1 += binding(q, site);
        std::printf("%.17g\n", q * (p2 - p1));
    }
    return 0;
}
