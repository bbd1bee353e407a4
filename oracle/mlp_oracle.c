/*
 * mlp_oracle.c -- FP64 forward and backward pass of the surrogate network.
 *
 * TEST INFRASTRUCTURE.  The parity checker for the GPU training and inference kernels
 * (BF16 operands, FP32 accumulation): the same network and loss in double precision.
 * The reference has no learned component, so there is no reference file to follow; the
 * network is
 *
 *     z1 = x W1 + b1,  a1 = silu(z1)         x: n_in,  W1: n_in x h
 *     z2 = a1 W2 + b2, a2 = silu(z2)         W2: h x h
 *     z3 = a2 W3 + b3, a3 = silu(z3)         W3: h x h
 *     y  = a3 . w4 + b4                      w4: h
 *     loss = (1/N) sum_n m_n (y_n - t_n)^2   (m_n = 0 masks a sample out)
 *
 * silu(z) = z / (1 + exp(-z)).  Parameters are one flat vector in the order W1 (row
 * major, input index outer), b1, W2, b2, W3, b3, w4, b4.  The backward pass is the chain
 * rule written out layer by layer; tests pin it with central finite differences.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

static double sig(double z) { return 1.0 / (1.0 + exp(-z)); }

long om_param_count(int n_in, int h) { return (long)n_in * h + h + 2L * ((long)h * h + h) + h + 1; }

typedef struct {
    const double *W1, *b1, *W2, *b2, *W3, *b3, *w4, *b4;
} om_view;

static om_view view(const double *p, int n_in, int h) {
    om_view v;
    v.W1 = p;
    v.b1 = v.W1 + (long)n_in * h;
    v.W2 = v.b1 + h;
    v.b2 = v.W2 + (long)h * h;
    v.W3 = v.b2 + h;
    v.b3 = v.W3 + (long)h * h;
    v.w4 = v.b3 + h;
    v.b4 = v.w4 + h;
    return v;
}

/* one dense layer with silu: z = a W + b, out = silu(z) */
static void layer(const double *a, int k, const double *W, const double *b, int h, double *z, double *out) {
    for (int j = 0; j < h; ++j) {
        double s = b[j];
        for (int i = 0; i < k; ++i) s += a[i] * W[(long)i * h + j];
        z[j] = s;
        out[j] = s * sig(s);
    }
}

/* predictions y[n] of rows x[n][n_in] */
void om_forward(const double *params, int n_in, int h, const double *x, int n, double *y) {
    om_view v = view(params, n_in, h);
    double *z = malloc(sizeof(double) * h), *a1 = malloc(sizeof(double) * h), *a2 = malloc(sizeof(double) * h),
           *a3 = malloc(sizeof(double) * h);
    for (int r = 0; r < n; ++r) {
        layer(x + (long)r * n_in, n_in, v.W1, v.b1, h, z, a1);
        layer(a1, h, v.W2, v.b2, h, z, a2);
        layer(a2, h, v.W3, v.b3, h, z, a3);
        double s = v.b4[0];
        for (int j = 0; j < h; ++j) s += a3[j] * v.w4[j];
        y[r] = s;
    }
    free(z); free(a1); free(a2); free(a3);
}

/* loss and its gradient (grad laid out as params); mask may be NULL (all ones) */
double om_loss_grad(const double *params, int n_in, int h, const double *x, const double *t, const double *mask, int n,
                    double *grad) {
    om_view v = view(params, n_in, h);
    const long np = om_param_count(n_in, h);
    memset(grad, 0, sizeof(double) * np);
    double *g = grad;
    double *gW1 = g, *gb1 = gW1 + (long)n_in * h, *gW2 = gb1 + h, *gb2 = gW2 + (long)h * h, *gW3 = gb2 + h,
           *gb3 = gW3 + (long)h * h, *gw4 = gb3 + h, *gb4 = gw4 + h;
    double *z1 = malloc(sizeof(double) * h), *z2 = malloc(sizeof(double) * h), *z3 = malloc(sizeof(double) * h);
    double *a1 = malloc(sizeof(double) * h), *a2 = malloc(sizeof(double) * h), *a3 = malloc(sizeof(double) * h);
    double *d3 = malloc(sizeof(double) * h), *d2 = malloc(sizeof(double) * h), *d1 = malloc(sizeof(double) * h);
    double loss = 0.0;
    for (int r = 0; r < n; ++r) {
        const double *xr = x + (long)r * n_in;
        const double m = mask ? mask[r] : 1.0;
        layer(xr, n_in, v.W1, v.b1, h, z1, a1);
        layer(a1, h, v.W2, v.b2, h, z2, a2);
        layer(a2, h, v.W3, v.b3, h, z3, a3);
        double y = v.b4[0];
        for (int j = 0; j < h; ++j) y += a3[j] * v.w4[j];
        const double e = y - t[r];
        loss += m * e * e / n;
        const double dy = 2.0 * m * e / n;
        gb4[0] += dy;
        for (int j = 0; j < h; ++j) {
            gw4[j] += dy * a3[j];
            const double s = sig(z3[j]);
            d3[j] = dy * v.w4[j] * s * (1.0 + z3[j] * (1.0 - s));
        }
        for (int i = 0; i < h; ++i) {
            double acc = 0.0;
            for (int j = 0; j < h; ++j) {
                gW3[(long)i * h + j] += a2[i] * d3[j];
                acc += d3[j] * v.W3[(long)i * h + j];
            }
            const double s = sig(z2[i]);
            d2[i] = acc * s * (1.0 + z2[i] * (1.0 - s));
        }
        for (int j = 0; j < h; ++j) gb3[j] += d3[j];
        for (int i = 0; i < h; ++i) {
            double acc = 0.0;
            for (int j = 0; j < h; ++j) {
                gW2[(long)i * h + j] += a1[i] * d2[j];
                acc += d2[j] * v.W2[(long)i * h + j];
            }
            const double s = sig(z1[i]);
            d1[i] = acc * s * (1.0 + z1[i] * (1.0 - s));
        }
        for (int j = 0; j < h; ++j) gb2[j] += d2[j];
        for (int i = 0; i < n_in; ++i)
            for (int j = 0; j < h; ++j) gW1[(long)i * h + j] += xr[i] * d1[j];
        for (int j = 0; j < h; ++j) gb1[j] += d1[j];
    }
    free(z1); free(z2); free(z3); free(a1); free(a2); free(a3); free(d1); free(d2); free(d3);
    return loss;
}
