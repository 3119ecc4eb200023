// FP64 MLP (forward, backward, Adam) and the radiance-cache input encoding.
// TEST INFRASTRUCTURE ONLY.  The radiance cache has no counterpart in the
// reference; this is the FP64 definition the BF16 GPU kernels are checked
// against.  Layout per linear layer: W (out x in, row-major) then b (out).
// Hidden layers use ReLU, the output layer exp.
#include <cmath>
#include <vector>

#include "oracle.h"

namespace {
struct Shape {
    uint32_t in, h, nh, out;
    uint32_t layers() const { return nh + 1; }
    uint32_t fan_in(uint32_t l) const { return l == 0 ? in : h; }
    uint32_t fan_out(uint32_t l) const { return l == nh ? out : h; }
};
size_t offset(const Shape& s, uint32_t l) {
    size_t o = 0;
    for (uint32_t k = 0; k < l; ++k) o += size_t(s.fan_in(k)) * s.fan_out(k) + s.fan_out(k);
    return o;
}
}  // namespace

extern "C" size_t or_mlp_param_count(uint32_t in, uint32_t h, uint32_t nh, uint32_t out) {
    Shape s{in, h, nh, out};
    return offset(s, s.layers());
}

namespace {
// acts[l] = input of layer l (post-activation of layer l-1); pre[l] = W x + b.
void forward(const double* p, const Shape& s, const double* x, std::vector<std::vector<double>>& acts,
             std::vector<std::vector<double>>& pre) {
    acts.assign(s.layers() + 1, {});
    pre.assign(s.layers(), {});
    acts[0].assign(x, x + s.in);
    for (uint32_t l = 0; l < s.layers(); ++l) {
        uint32_t fi = s.fan_in(l), fo = s.fan_out(l);
        const double* W = p + offset(s, l);
        const double* b = W + size_t(fi) * fo;
        pre[l].assign(fo, 0.0);
        acts[l + 1].assign(fo, 0.0);
        for (uint32_t o = 0; o < fo; ++o) {
            double z = b[o];
            for (uint32_t i = 0; i < fi; ++i) z += W[size_t(o) * fi + i] * acts[l][i];
            pre[l][o] = z;
            acts[l + 1][o] = (l == s.nh) ? std::exp(z) : (z > 0.0 ? z : 0.0);
        }
    }
}
}  // namespace

extern "C" void or_mlp_forward(const double* p, uint32_t in, uint32_t h, uint32_t nh, uint32_t out,
                               const double* x, size_t batch, double* y) {
    Shape s{in, h, nh, out};
    std::vector<std::vector<double>> acts, pre;
    for (size_t n = 0; n < batch; ++n) {
        forward(p, s, x + n * in, acts, pre);
        for (uint32_t c = 0; c < out; ++c) y[n * out + c] = acts[s.layers()][c];
    }
}

extern "C" double or_mlp_loss_grad(const double* p, uint32_t in, uint32_t h, uint32_t nh, uint32_t out,
                                   const double* x, const double* target, size_t batch, int rel, double eps,
                                   double* grad) {
    Shape s{in, h, nh, out};
    size_t np = offset(s, s.layers());
    for (size_t i = 0; i < np; ++i) grad[i] = 0.0;
    std::vector<std::vector<double>> acts, pre;
    double loss = 0.0;
    const double norm = 1.0 / (double(batch) * out);
    for (size_t n = 0; n < batch; ++n) {
        forward(p, s, x + n * in, acts, pre);
        const std::vector<double>& y = acts[s.layers()];
        std::vector<double> g(out);
        for (uint32_t c = 0; c < out; ++c) {
            double d = y[c] - target[n * out + c];
            double den = rel ? (y[c] * y[c] + eps) : 1.0;  // stop-gradient denominator
            loss += d * d / den * norm;
            g[c] = 2.0 * d / den * norm * y[c];  // through exp
        }
        for (int l = int(s.layers()) - 1; l >= 0; --l) {
            uint32_t fi = s.fan_in(l), fo = s.fan_out(l);
            const double* W = p + offset(s, l);
            double* gW = grad + offset(s, l);
            double* gb = gW + size_t(fi) * fo;
            std::vector<double> gin(fi, 0.0);
            for (uint32_t o = 0; o < fo; ++o) {
                gb[o] += g[o];
                for (uint32_t i = 0; i < fi; ++i) {
                    gW[size_t(o) * fi + i] += g[o] * acts[l][i];
                    gin[i] += W[size_t(o) * fi + i] * g[o];
                }
            }
            if (l > 0)
                for (uint32_t i = 0; i < fi; ++i) gin[i] = pre[l - 1][i] > 0.0 ? gin[i] : 0.0;
            g.swap(gin);
        }
    }
    return loss;
}

extern "C" void or_adam_step(double* p, double* m, double* v, const double* g, size_t n, double lr, double b1,
                             double b2, double eps, uint32_t t) {
    double c1 = 1.0 - std::pow(b1, double(t));
    double c2 = 1.0 - std::pow(b2, double(t));
    for (size_t i = 0; i < n; ++i) {
        m[i] = b1 * m[i] + (1.0 - b1) * g[i];
        v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
        double mh = m[i] / c1, vh = v[i] / c2;
        p[i] -= lr * mh / (std::sqrt(vh) + eps);
    }
}

extern "C" void or_cache_encode(const double pos01[3], const double dir[3], const double nrm[3],
                                const double albedo[3], double out[81]) {
    for (int a = 0; a < 3; ++a) {
        for (int f = 0; f < 12; ++f) {
            double arg = std::ldexp(M_PI, f) * pos01[a];
            out[a * 24 + 2 * f] = std::sin(arg);
            out[a * 24 + 2 * f + 1] = std::cos(arg);
        }
    }
    for (int k = 0; k < 3; ++k) {
        out[72 + k] = dir[k];
        out[75 + k] = nrm[k];
        out[78 + k] = albedo[k];
    }
}
