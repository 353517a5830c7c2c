// ref_eld_harness.cpp -- links the reference's own src/ElasticMeasures.cpp
// (compiled from /root/reference by oracle/Makefile into oracle/_ref/) and
// exposes its verifiers through a C ABI for the oracle's parity tests.
//
// TEST INFRASTRUCTURE ONLY.  ElasticMeasures.cpp is the one reference
// translation unit that builds from its own source: every other one includes
// include/funcs.h, which needs FFTW3's fftw3.h (absent here).
//
// The prototypes below match include/funcs.h:76-82 of the reference; the
// dispatch follows src/fsmeld.cpp:170-183 (argument order per metric).
extern double dtw(double* A, int m, double* B, int n, int r);
extern double lcss(double* q, int n, double* A, int m, int w, double epsilon);
extern double erp(double* x, int lenx, double* y, int leny, double g, int w);
extern double edr(double* x, int n, double* y, int m_size, double epsilon, int w);
extern double msm(double* x, int xlen, double* y, int ylen, double c, int w);
extern double swale(double* x, int m, double* y, int n, double p, double r_cost, double epsilon,
                    int w);

extern "C" double ref_eld(int metric, double* q, double* t, int m, int w, double arg) {
    switch (metric) {
        case 5: return msm(q, m, t, m, arg, w);
        case 3: return erp(q, m, t, m, arg, w);
        case 4: return lcss(q, m, t, m, w, arg);
        case 2: return edr(q, m, t, m, arg, w);
        case 1: return dtw(q, m, t, m, w);
        // SWALE with the costs of include/funcs.h:30-31 (mismatch 5, match 1)
        case 6: return swale(q, m, t, m, 5.0, 1.0, arg, w);
        default: return __builtin_nan("");
    }
}
