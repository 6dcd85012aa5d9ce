/*
 * ORACLE — test infrastructure only.  Nothing in the product links this.
 *
 * Timing harness of the CPU baseline: the BSIM3 Master loops of the
 * oracle (the reference's MOSFET_B3::Master::updateState /
 * loadDAEVectors / loadDAEMatrices, N_DEV_MOSFET_B3.C:9222-9928) run
 * serially, or partitioned over host threads the way Xyce distributes
 * devices over MPI ranks: each partition owns a contiguous block of
 * instances and loads into its own copy of the vectors and matrices
 * (the overlap vectors of a rank), which are then added into the owned
 * result in fixed partition order (the Export(Add) of fillComplete,
 * N_LAS_Matrix.C:197-204; N_DEV_DeviceMgr.C:2230-2232, 2277-2282).  The
 * sum is deterministic run to run; it rounds differently from the serial
 * load only where two partitions add into the same entry.
 */
#include <pthread.h>
#include <stdint.h>
#include <string.h>
#include <time.h>

typedef struct ref_b3_master ref_b3_master;
int ref_b3_update_state(ref_b3_master *M, double *nextSol, double *currSta,
                        double *currSto, double *nextSta, double *nextSto);
void ref_b3_load_vectors(ref_b3_master *M, double *fVec, double *qVec,
                         double *fl, double *ql);
void ref_b3_load_matrices(ref_b3_master *M);

typedef struct {
  int P;
  ref_b3_master **M;
  double *x, *sta_c, *sto_c, *sta_n, *sto_n;
  double **part;   /* P x 6 private outputs: F, Q, FL, QL (n1), JF, JQ (nnz) */
  double **out;    /* 6 combined outputs */
  int64_t n1, nnz;
  int reps;
  pthread_barrier_t bar;
  /* start gate: workers wait until every thread exists (go) or creation
     failed (abort), so none waits at a barrier that cannot fill */
  pthread_mutex_t mu;
  pthread_cond_t cv;
  int go, abort_;
} job_t;

typedef struct {
  job_t *j;
  int t;
} arg_t;

static void *worker(void *p) {
  arg_t *a = (arg_t *)p;
  job_t *j = a->j;
  const int t = a->t, P = j->P;
  int r, k, q;
  double **mine = j->part + (size_t)6 * t;
  pthread_mutex_lock(&j->mu);
  while (!j->go && !j->abort_) pthread_cond_wait(&j->cv, &j->mu);
  pthread_mutex_unlock(&j->mu);
  if (j->abort_) return NULL;
  for (r = 0; r < j->reps; ++r) {
    /* the assembler zeroes the outputs, then the three Master calls */
    for (k = 0; k < 6; ++k)
      memset(mine[k], 0, sizeof(double) * (size_t)(k < 4 ? j->n1 : j->nnz));
    ref_b3_update_state(j->M[t], j->x, j->sta_c, j->sto_c, j->sta_n, j->sto_n);
    ref_b3_load_vectors(j->M[t], mine[0], mine[1], mine[2], mine[3]);
    ref_b3_load_matrices(j->M[t]);
    if (P == 1) continue;
    pthread_barrier_wait(&j->bar);
    /* fixed-order combine, every thread over its share of the entries */
    for (k = 0; k < 6; ++k) {
      const int64_t n = k < 4 ? j->n1 : j->nnz;
      const int64_t lo = n * t / P, hi = n * (t + 1) / P;
      double *o = j->out[k];
      int64_t i;
      for (i = lo; i < hi; ++i) {
        double s = j->part[k][i];
        for (q = 1; q < P; ++q) s += j->part[(size_t)6 * q + k][i];
        o[i] = s;
      }
    }
    pthread_barrier_wait(&j->bar);
  }
  return NULL;
}

/* Runs `reps` full loads with P partitions (P masters, each set up over its
   own instances with its matrix pointers into part[6 t + 4 / 5]); returns
   the wall time in seconds, or -1 if threads could not be started.  With
   P = 1 the single master's outputs are part[0..5] and nothing is
   combined. */
double ref_b3_partitioned_loads(int P, ref_b3_master **M, double *x, double *sta_c,
                                double *sto_c, double *sta_n, double *sto_n,
                                double **part, double **out, int64_t n1, int64_t nnz,
                                int reps) {
  job_t j;
  pthread_t th[256];
  arg_t args[256];
  struct timespec t0, t1;
  int t, started = 0;
  if (P < 1 || P > 256) return -1.0;
  j.P = P; j.M = M; j.x = x; j.sta_c = sta_c; j.sto_c = sto_c;
  j.sta_n = sta_n; j.sto_n = sto_n; j.part = part; j.out = out;
  j.n1 = n1; j.nnz = nnz; j.reps = reps;
  j.go = 0;
  j.abort_ = 0;
  pthread_mutex_init(&j.mu, NULL);
  pthread_cond_init(&j.cv, NULL);
  if (P > 1 && pthread_barrier_init(&j.bar, NULL, (unsigned)P) != 0) return -1.0;
  for (t = 1; t < P; ++t) {
    args[t].j = &j;
    args[t].t = t;
    if (pthread_create(&th[t], NULL, worker, &args[t]) != 0) break;
    started++;
  }
  pthread_mutex_lock(&j.mu);
  if (started == P - 1) j.go = 1;
  else j.abort_ = 1;
  pthread_cond_broadcast(&j.cv);
  pthread_mutex_unlock(&j.mu);
  clock_gettime(CLOCK_MONOTONIC, &t0);
  if (j.go) {
    args[0].j = &j;
    args[0].t = 0;
    worker(&args[0]);
  }
  for (t = 1; t <= started; ++t) pthread_join(th[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (P > 1) pthread_barrier_destroy(&j.bar);
  pthread_cond_destroy(&j.cv);
  pthread_mutex_destroy(&j.mu);
  if (!j.go) return -1.0;
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
