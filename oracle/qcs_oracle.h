/* qcs_oracle.h -- CPU FP64 oracle (test infrastructure; see qcs_oracle.c). */
#ifndef QCS_ORACLE_H
#define QCS_ORACLE_H

/* op kinds of oq_run / oq_gate */
#define OQ_H 0
#define OQ_X 1
#define OQ_Y 2
#define OQ_Z 3
#define OQ_P 4
#define OQ_RX 5
#define OQ_RY 6
#define OQ_RZ 7
#define OQ_CNOT 8
#define OQ_CPHASE 9
#define OQ_SWAP 10

typedef struct oq_state oq_state;

oq_state *oq_create(int n, int compat);
void oq_destroy(oq_state *s);
long oq_size(const oq_state *s);
double *oq_vector(oq_state *s);  /* 2*size doubles {re, im} */
double *oq_scratch(oq_state *s);

void oq_apply_1q(oq_state *s, int t, const double m[8]);
void oq_apply_cu(oq_state *s, int c, int t, const double m[8]);
void oq_apply_swap(oq_state *s, int a, int b);
void oq_phase_flip(oq_state *s, long index);
void oq_diffusion(oq_state *s);
long oq_grover_iterations(int n, int compat);
double oq_normalize(oq_state *s);
double oq_norm_sq(const oq_state *s);
int oq_measure(oq_state *s, int q, double u);
void oq_sample(const oq_state *s, int shots, const double *u, long *out);
double oq_probability(const oq_state *s, long i);
long oq_argmax(const oq_state *s);
void oq_gate(int kind, double angle, double m[8]);
void oq_run(oq_state *s, int ng, const int *op, const int *tq, const int *cq,
            const double *par);
void oq_qft(oq_state *s);
void oq_grover(oq_state *s, long marked, long iterations);
void oq_apply_1q_omp(oq_state *s, int t, const double m[8], int racy);
void oq_apply_cu_omp(oq_state *s, int c, int t, const double m[8], int mode);
void oq_run_omp(oq_state *s, int ng, const int *op, const int *tq, const int *cq,
                const double *par, int racy);
void oq_set_threads(int threads);

#endif
