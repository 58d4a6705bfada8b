// Host-side sanitizer driver (test infrastructure): runs xsh_plan - the
// library's descriptor validation, lattice derivation, kernel choice and
// every host-built table - over descriptors of every lattice family, boundary
// type, table kind and refusal path.  Built with AddressSanitizer and
// UndefinedBehaviorSanitizer on the host code of the C-ABI unit; needs no
// device.  Exits non-zero when a case returns an unexpected status (the
// sanitizers abort on their own findings).
#include <complex>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/xspde_hip.h"

namespace {

xsh_desc base(int prec, int dims, int nx, int ny) {
  xsh_desc d;
  std::memset(&d, 0, sizeof d);
  d.struct_size = (int)sizeof d;
  d.precision = prec;
  d.dimensions = dims;
  d.points = 11;
  d.nx = dims > 1 ? nx : 1;
  d.ny = dims > 2 ? ny : 1;
  d.steps = 2;
  d.checks = 1;
  d.method = XSH_RK4;
  d.iterations = 4;
  d.ensembles[0] = 33;
  d.ensembles[1] = 3;
  d.ensembles[2] = 2;
  d.noises = 2;
  d.inrandoms = 2;
  d.ranges[0] = 1.0;
  d.ranges[1] = 20.0;
  d.ranges[2] = 15.0;
  d.origins[1] = -10.0;
  d.origins[2] = -7.5;
  d.deriv = XSH_DERIV_GP;
  d.deriv_par[0] = 0.0;
  d.deriv_par[1] = -1.0;
  d.deriv_par[2] = -0.1;
  d.deriv_par[4] = 0.2;
  d.has_linear = 1;
  d.lin_l0[0] = 0.5;
  d.lin_l2[0][1] = 0.5;
  d.lin_l2[1][1] = 0.5;
  d.lin_l1[0][0] = 0.1;
  d.initial = XSH_INIT_SECH;
  d.init_par[0] = 1.0;
  d.init_par[1] = 1.0;
  d.init_par[2] = 0.1;
  d.nobserve = 3;
  d.observe[0] = XSH_OBS_ABS2;
  d.observe[1] = XSH_OBS_A_AVE;
  d.observe[2] = XSH_OBS_ABS2_INT;
  return d;
}

int failures = 0, cases = 0;

void expect(const char* what, const xsh_desc& d, int want) {
  xsh_info info;
  const int rc = xsh_plan(&d, &info);
  ++cases;
  if (rc != want) {
    ++failures;
    std::printf("FAIL %-44s rc %d (want %d): %s\n", what, rc, want, xsh_last_error());
  } else if (rc == XSH_OK) {
    std::printf("ok   %-44s kind %d threads %4d tpb %3d avlen %lld\n", what, info.kernel_kind, info.threads,
                info.traj_per_block, info.avlen);
  } else {
    std::printf("ok   %-44s refused: %s\n", what, xsh_last_error());
  }
}

}  // namespace

int main() {
  const int precs[2] = {XSH_FP32, XSH_FP64};
  for (int prec : precs) {
    const char* pn = prec == XSH_FP64 ? "fp64" : "fp32";
    char name[96];
    // SDE, with a time-domain spectrum
    xsh_desc d = base(prec, 1, 1, 1);
    d.tspec[1] = 1;
    std::snprintf(name, sizeof name, "%s SDE + spectrum", pn);
    expect(name, d, XSH_OK);
    // 1D: every FFT size, every boundary type, padded sizes
    for (int n : {32, 35, 64, 101, 128, 161, 256, 512, 1024}) {
      for (int t = 0; t < 5; ++t) {
        static const int b[5][2] = {{0, 0}, {-1, -1}, {1, 1}, {-1, 1}, {1, -1}};
        d = base(prec, 2, n, 1);
        d.boundaries[0][0] = b[t][0];
        d.boundaries[0][1] = b[t][1];
        d.boundval[0][1][0] = t ? 0.3 : 0.0;
        const bool ok = t == 0 ? true : n <= 256;
        std::snprintf(name, sizeof name, "%s 1D %d boundary [%d,%d]", pn, n, b[t][0], b[t][1]);
        expect(name, d, ok ? XSH_OK : XSH_EUNSUPPORTED);
      }
    }
    // 2D: FFT sizes, dense / padded / mixed lattices, the streaming kernels
    const int shapes[][2] = {{32, 32}, {64, 64}, {128, 128}, {35, 35}, {48, 40}, {70, 34}, {128, 64}, {64, 128}, {33, 64}};
    for (auto& sh : shapes) {
      for (int t = 0; t < 3; ++t) {
        d = base(prec, 3, sh[0], sh[1]);
        if (t == 1) d.boundaries[0][0] = d.boundaries[0][1] = d.boundaries[1][0] = d.boundaries[1][1] = -1;
        if (t == 2) d.boundaries[1][0] = d.boundaries[1][1] = 1;
        std::snprintf(name, sizeof name, "%s 2D %dx%d type %d", pn, sh[0], sh[1], t);
        xsh_info info;
        const int rc = xsh_plan(&d, &info);  // some families refuse some combinations: report, not fail
        ++cases;
        std::printf("%s %-44s rc %d%s%s\n", rc == XSH_OK || rc == XSH_EUNSUPPORTED ? "ok  " : "FAIL", name, rc,
                    rc ? ": " : "", rc ? xsh_last_error() : "");
        if (rc != XSH_OK && rc != XSH_EUNSUPPORTED) ++failures;
      }
    }
    // a boundfun table on a padded 2D lattice (one level per half fine step)
    d = base(prec, 3, 48, 40);
    d.boundaries[0][0] = d.boundaries[0][1] = 1;
    const int calls = d.points * d.steps - 1;
    d.boundlevels = 2 * calls * (1 + d.checks) + 1;
    std::vector<double> tab((size_t)d.boundlevels * 2 * 2 * 48 * 2);
    for (size_t i = 0; i < tab.size(); ++i) tab[i] = 1e-3 * (double)(i % 97);
    d.boundtab = tab.data();
    std::snprintf(name, sizeof name, "%s 48x40 Dirichlet boundfun table", pn);
    expect(name, d, XSH_OK);
    // k-space filtered noises and randoms
    d = base(prec, 3, 64, 64);
    d.noises = 0;
    d.inrandoms = 0;
    d.knoises = 2;
    d.krandoms = 2;
    std::vector<double> nf(2 * 64 * 64 * 2, 0.5), rf(2 * 64 * 64 * 2, 0.25);
    d.nfilter = nf.data();
    d.rfilter = rf.data();
    std::snprintf(name, sizeof name, "%s 64x64 k-space noises", pn);
    expect(name, d, XSH_OK);
    // truncated-Wigner features
    d = base(prec, 3, 128, 128);
    d.initial = XSH_INIT_COHERENT;
    d.init_par[0] = 10.0;
    d.checks = 0;
    d.nobserve = 3;
    d.observe[0] = XSH_OBS_ABS2_AVE;
    d.observe[1] = XSH_OBS_ABS4_AVE;
    d.observe[2] = XSH_OBS_LOWK;
    d.traj_out = 1;
    d.nmodes = 4;
    const int modes[4][2] = {{0, 0}, {1, 0}, {0, 1}, {127, 0}};
    std::memcpy(d.modes, modes, sizeof modes);
    std::snprintf(name, sizeof name, "%s 128x128 TW features", pn);
    expect(name, d, prec == XSH_FP32 ? XSH_OK : XSH_EUNSUPPORTED);
    // coupled SDE components (the matrix drift form), 2..4 components
    for (int F = 2; F <= XSH_MAX_FIELDS; ++F) {
      d = base(prec, 1, 1, 1);
      d.fields = F;
      d.deriv = XSH_DERIV_MATRIX;
      d.initial = XSH_INIT_CONST;
      d.noises = F;
      for (int i = 0; i < F; ++i) {
        d.fmat[i][(i + 1) % F][0] = 1.0;
        d.fmat[(i + 1) % F][i][0] = -1.0;
        d.fnl[i][1] = 0.1;
        d.fnoise[i][i % 4] = 0.2;
        d.fl0[i][0] = -0.1;
        d.fc0[i][0] = 1.0 + i;
      }
      d.nobserve = 3;
      d.observe[0] = XSH_OBS_A;
      d.observe[1] = XSH_OBS_ABS2_AVE;
      d.observe[2] = XSH_OBS_ABS4_AVE;
      d.obs_field[0] = F - 1;
      d.obs_field[1] = 0;
      d.obs_field[2] = 1;
      std::snprintf(name, sizeof name, "%s SDE %d coupled components", pn, F);
      expect(name, d, XSH_OK);
    }
    d = base(prec, 1, 1, 1);
    d.fields = 2;
    d.deriv = XSH_DERIV_MATRIX;
    d.initial = XSH_INIT_CONST;
    d.obs_field[0] = 2;
    expect("coupled: observable of a missing component", d, XSH_EINVAL);
    d.obs_field[0] = 0;
    d.observe[0] = XSH_OBS_CORR;
    expect("coupled: lattice observable", d, XSH_EUNSUPPORTED);
    d = base(prec, 3, 64, 64);
    d.fields = 2;
    d.deriv = XSH_DERIV_MATRIX;
    expect("coupled lattice components", d, XSH_EUNSUPPORTED);
    d = base(prec, 1, 1, 1);
    d.deriv = XSH_DERIV_MATRIX;
    expect("matrix form with one component", d, XSH_EINVAL);
    // refusals
    d = base(prec, 3, 64, 64);
    d.tspec[0] = 1;
    expect("spectrum on a lattice", d, XSH_EUNSUPPORTED);
    d = base(prec, 2, 64, 1);
    d.boundaries[0][0] = 2;
    expect("invalid boundary", d, XSH_EINVAL);
    d = base(prec, 2, 64, 1);
    d.ensembles[1] = 0;
    expect("ensembles(2) = 0", d, XSH_EINVAL);
    d = base(prec, 2, 64, 1);
    d.struct_size = 8;
    expect("descriptor size", d, XSH_EINVAL);
    d = base(prec, 3, 129, 64);
    expect("odd 2D lattice above 128", d, XSH_EUNSUPPORTED);
  }
  std::printf("%d cases, %d failures\n", cases, failures);
  return failures ? 1 : 0;
}
