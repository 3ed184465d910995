// TEST INFRASTRUCTURE ONLY: binary64 restatement of the learned blend
// surrogate (forward, loss and gradients), the checker of the GPU training
// kernels.  The network has no counterpart in the reference; its
// definition is the one in include/i360hip.h (6 layers of 3x3 convolution,
// 16 -> 32 x5 -> 4 channels, ReLU between layers, zero padding; loss
//   L = mean_p [ 1e4 * mean_c d_c^2 + 1e-2 * (o_3 - log10(mean_c d_c^2 + 1e-10))^2 ],
//   d = o_{0..2}/64 - (target - feat_{7..9}),
// with the error term's target held constant).
//
// Weights arrive as float values (the caller rounds them to BF16 first, as
// the GPU uses them); everything else is computed in binary64.
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

constexpr int kLayers = 6;
int cin_of(int j) { return j == 0 ? 16 : 32; }
int cout_of(int j) { return j == kLayers - 1 ? 4 : 32; }
int woff(int j) {
    int o = 0;
    for (int k = 0; k < j; ++k) o += 9 * cin_of(k) * cout_of(k);
    return o;
}
int boff(int j) {
    int o = 0;
    for (int k = 0; k < j; ++k) o += cout_of(k);
    return o;
}

// z = conv3x3(x) + b over n images of h x w, NHWC, zero padding.
void conv(const std::vector<double>& x, int cin, const float* w, const float* b, int cout, int n, int h, int wd,
          std::vector<double>& z) {
    z.assign(static_cast<size_t>(n) * h * wd * cout, 0.0);
    for (int i = 0; i < n; ++i)
        for (int y = 0; y < h; ++y)
            for (int xx = 0; xx < wd; ++xx) {
                double* o = &z[((static_cast<size_t>(i) * h + y) * wd + xx) * cout];
                for (int co = 0; co < cout; ++co) o[co] = b ? b[co] : 0.0;
                for (int k = 0; k < 9; ++k) {
                    const int sy = y + k / 3 - 1, sx = xx + k % 3 - 1;
                    if (sy < 0 || sx < 0 || sy >= h || sx >= wd) continue;
                    const double* in = &x[((static_cast<size_t>(i) * h + sy) * wd + sx) * cin];
                    for (int ci = 0; ci < cin; ++ci) {
                        const float* wr = w + (static_cast<size_t>(k) * cin + ci) * cout;
                        for (int co = 0; co < cout; ++co) o[co] += static_cast<double>(wr[co]) * in[ci];
                    }
                }
            }
}

void forward(const float* w, const float* b, const float* feat, int n, int h, int wd,
             std::vector<std::vector<double>>& act) {
    std::vector<double> x(feat, feat + static_cast<size_t>(n) * h * wd * 16);
    act.assign(kLayers, {});
    for (int j = 0; j < kLayers; ++j) {
        const std::vector<double>& in = j == 0 ? x : act[j - 1];
        conv(in, cin_of(j), w + woff(j), b + boff(j), cout_of(j), n, h, wd, act[j]);
        if (j < kLayers - 1)
            for (double& v : act[j]) v = v > 0.0 ? v : 0.0;
    }
}

}  // namespace

extern "C" {

void orc_surr_forward(const float* w, const float* b, const float* feat, int n, int h, int wd, float* out) {
    std::vector<std::vector<double>> act;
    forward(w, b, feat, n, h, wd, act);
    const std::vector<double>& o = act[kLayers - 1];
    for (size_t i = 0; i < o.size(); ++i) out[i] = static_cast<float>(o[i]);
}

// le_in (optional, one value per pixel): use it as the error term's target
// instead of log10(mean_c d_c^2 + 1e-10) (the loss then is the function
// whose gradient is returned, for finite-difference checks); le_out
// (optional) receives the computed log10 values.
void orc_surr_loss_grad(const float* w, const float* b, const float* feat, const float* target, int n, int h, int wd,
                        double* gw, double* gb, double* loss, const double* le_in, double* le_out) {
    std::vector<std::vector<double>> act;
    forward(w, b, feat, n, h, wd, act);
    const size_t P = static_cast<size_t>(n) * h * wd;
    const std::vector<double>& o = act[kLayers - 1];
    std::vector<double> g(P * 4);
    double L = 0.0;
    for (size_t p = 0; p < P; ++p) {
        double d[3], e2 = 0.0;
        for (int c = 0; c < 3; ++c) {
            d[c] = o[4 * p + c] / 64.0 - (static_cast<double>(target[3 * p + c]) - feat[16 * p + 7 + c]);
            e2 += d[c] * d[c];
        }
        e2 /= 3.0;
        const double le = le_in ? le_in[p] : std::log10(e2 + 1e-10);
        if (le_out) le_out[p] = le;
        const double de = o[4 * p + 3] - le;
        L += 1e4 * e2 + 1e-2 * de * de;
        for (int c = 0; c < 3; ++c) g[4 * p + c] = 1e4 * (2.0 / 3.0) * d[c] / 64.0 / static_cast<double>(P);
        g[4 * p + 3] = 1e-2 * 2.0 * de / static_cast<double>(P);
    }
    *loss = L / static_cast<double>(P);
    for (int i = 0; i < woff(kLayers); ++i) gw[i] = 0.0;
    for (int i = 0; i < boff(kLayers); ++i) gb[i] = 0.0;
    std::vector<double> x(feat, feat + P * 16);
    for (int j = kLayers - 1; j >= 0; --j) {
        const int ci_n = cin_of(j), co_n = cout_of(j);
        // g holds dL/d(output of layer j); through the ReLU
        if (j < kLayers - 1)
            for (size_t i = 0; i < g.size(); ++i)
                if (!(act[j][i] > 0.0)) g[i] = 0.0;
        const std::vector<double>& in = j == 0 ? x : act[j - 1];
        double* gwj = gw + woff(j);
        double* gbj = gb + boff(j);
        std::vector<double> gin(P * ci_n, 0.0);
        const float* wj = w + woff(j);
        for (int i = 0; i < n; ++i)
            for (int y = 0; y < h; ++y)
                for (int xx = 0; xx < wd; ++xx) {
                    const size_t p = (static_cast<size_t>(i) * h + y) * wd + xx;
                    const double* gz = &g[p * co_n];
                    for (int co = 0; co < co_n; ++co) gbj[co] += gz[co];
                    for (int k = 0; k < 9; ++k) {
                        const int sy = y + k / 3 - 1, sx = xx + k % 3 - 1;
                        if (sy < 0 || sx < 0 || sy >= h || sx >= wd) continue;
                        const size_t q = (static_cast<size_t>(i) * h + sy) * wd + sx;
                        for (int ci = 0; ci < ci_n; ++ci) {
                            const size_t wi = (static_cast<size_t>(k) * ci_n + ci) * co_n;
                            double acc = 0.0;
                            for (int co = 0; co < co_n; ++co) {
                                gwj[wi + co] += gz[co] * in[q * ci_n + ci];
                                acc += static_cast<double>(wj[wi + co]) * gz[co];
                            }
                            gin[q * ci_n + ci] += acc;
                        }
                    }
                }
        g.swap(gin);
    }
}

}  // extern "C"
