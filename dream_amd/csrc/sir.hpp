// sir.hpp -- host side of the SIR (dreamrect / dream_arr_rect) kernels.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace dream {

// One dream_arr_rect problem as the host sees it (host copies of the
// per-element inputs; device pointers for Ro, delay and the output).
struct SirProblem {
  const double *Ro = nullptr;  int64_t No = 0;          // device
  int shape = 0;                                       // SIR_RECT, SIR_CIRC or SIR_CYLIND
  // Rectangles and cylinders: sides a (x), b (y); discs: radius R;
  // cylinders: curvature radius R (> 0 focused, < 0 defocused).
  double a = 0, b = 0, R = 0, dx = 0, dy = 0, dt = 0;  int64_t nt = 0;
  int delay_type = 0;  const double *delay = nullptr;   // device
  double v = 1, cp = 1000;
  std::vector<double> G;                               // L x 3 column-major (host)
  int foc_met = 0;  std::vector<double> focal;         // 1 or L values (host)
  int steer_met = 0;  double theta = 0, phi = 0;
  int do_apod = 0, apod_met = 0;  double apod_par = 0;
  std::vector<double> apod;                            // L values (host)
  int err_level = L_STOP;
  void *h = nullptr;                                   // device, nt x No
};

// Device descriptor of one problem (lives in device memory).
struct SirDev {
  const double *Ro;  int64_t No;
  const double *delay;  int32_t delay_multi;
  int32_t L;
  const double *xs;  const double *ys;  int32_t nx, ny, row_stride, chunks;
  // Discs: per-row first x position and sub-element count (doubles).
  const double *row_x0;  const double *row_n;  double dx;  int32_t shape, pad1;
  // Cylinders: per-row z offset of the curved surface (rz = z - zrow).
  const double *zrow;
  int32_t step_e, step_j, step_row, pad0;   // row walk: q += 256 as (e, j, row) steps
  const double *el;                 // 6 x L: gx gy gz foc steer num
  int64_t nt;
  double cp, dt, inv_dt, c1, amp;   // c1 = 1000/(cp dt), amp = 1000/(2 pi dt)
  void *h;
  Status *status;
  int64_t item_begin;               // first work item of this problem
  int32_t mode;                     // SIR_* below
  // Column walk (rectangles): the (element, x column) pairs, each cut into
  // nseg segments of seg_len rows, as units u = sg*(L*nx) + i*L + e; u +=
  // 256 steps (e, i, sg) by cstep_*.  0 segments: rows only; walk 1: every
  // point walks columns (else each point picks its cheaper walk).
  int32_t nseg;
  int32_t seg_len, cstep_e, cstep_i, cstep_sg, walk;
  // The walk choice: mean |xs| and |ys|, and dx, dy in samples of path
  // (dx / (cp dt/1000), dy / (cp dt/1000)).
  double wx, wy, sx, sy;
};

enum { SIR_LDS_STORE = 0, SIR_LDS_ATOMIC = 1, SIR_GLOBAL = 2 };
enum { SIR_RECT = 0, SIR_CIRC = 1, SIR_CYLIND = 2 };

// Sub-element rows of a disc of radius R (src/dreamcirc.cc:300-348): ys =
// -R + dy/2 while ys <= R; in row ys the x walk starts at -sqrt(R*R - ys*ys)
// + dx/2 and runs while xs <= sqrt(R*R - ys*ys).  Fills ys, the first x and
// the count of each row.
void disc_rows(double R, double dx, double dy, std::vector<double> &ys,
               std::vector<double> &x0, std::vector<double> &count);

// Cuts disc rows into segments of equal-ish length for lane balance
// (tables as disc_rows gives them; L elements per point).
void split_disc_rows(int64_t L, double dx, std::vector<double> &ys, std::vector<double> &x0,
                     std::vector<double> &count);

// Sub-element rows of a cylindrically curved element of y-size b and
// curvature radius Rc (src/dreamcylind.cc:288-350): the rows are angles phi
// from -phi_max + dphi/2 in steps dphi = asin(dy/|Rc|) while phi <= phi_max,
// phi_max = atan((b/2)/(|Rc| - z_Rc)), z_Rc = |Rc| - sqrt(Rc^2 - (b/2)^2);
// row k sits at ys = |Rc| sin(phi) and at depth zs = |Rc| - sqrt(Rc^2 - ys^2),
// stored as zrow = zs (focused, Rc >= 0: rz = z - zs) or -zs (defocused:
// rz = z + zs).  Returns dphi (the surface element is |Rc| dx dphi).
double cylinder_rows(double b, double Rc, double dy, std::vector<double> &ys,
                     std::vector<double> &zrow);

// Sub-element centres of one side: s = -w/2 + d/2; while (s <= w/2) s += d
// (the accumulated sequence of src/dreamrect.cc:286-331).
std::vector<double> subgrid(double w, double d);

// Per-element (gx gy gz foc steer num) table of one problem, evaluated on
// the host in the order of src/dream_arr_rect.cc:207-234.
std::vector<double> element_table(const SirProblem &p);

// A prepared batch: device tables, work list, status blocks, launch shape.
class SirPlan {
public:
  SirPlan(const std::vector<SirProblem> &problems, int out_dtype, int target_items = 0);
  ~SirPlan();
  void run(hipStream_t s, bool use_graph);
  // Per-problem SIRError code and OOB counts (synchronises on the last run's stream).
  void status(int32_t *codes, int64_t *oob);
  int64_t n_items() const { return n_items_; }

private:
  struct SirShape {
    bool persistent;   // grid pulls items from the counter
    int grid;
  };
  void choose(hipStream_t s);               // launch shape: autotuned or heuristic
  // With a gate, chunks of about chunk_work() contributions with
  // cancellation tests between them; without, one launch.
  void launch(hipStream_t s, ChunkGate *gate);
  void launch_shape(const SirShape &sh, hipStream_t s, ChunkGate *gate);
  std::vector<SirShape> shapes_;            // candidates, heuristic first
  int pick_ = -1;
  int out_dtype_;
  std::vector<SirDev> host_;                // problems in plan order
  std::vector<int64_t> item_count_;         // per problem (plan order): work items
  std::vector<double> item_cost_;           //   and contributions per item
  std::vector<int> levels_;
  std::vector<int64_t> zero_len_;  // per problem: elements to zero first (0 = none)
  std::vector<void *> zero_ptr_;
  SirDev *d_probs_ = nullptr;
  double *d_tables_ = nullptr;
  Status *d_status_ = nullptr;
  unsigned long long *d_counter_ = nullptr;
  int64_t n_items_ = 0;
  size_t lds_bytes_ = 0;
  int grid_ = 0;
  bool persistent_ = false;
  hipGraphExec_t graph_ = nullptr;
  hipStream_t last_stream_ = nullptr;
  bool ran_ = false;
};

// SIR with attenuation (alpha > DBL_EPSILON, in dB/(cm MHz)) of a list of
// problems, each written to its own h (overwritten); per-problem SIRError
// codes and out-of-bounds counts as SirPlan::status reports them.
// Synchronises `s`.
void run_sir_att(const std::vector<SirProblem> &problems, double alpha, int out_dtype,
                 hipStream_t s, int32_t *codes, int64_t *oobs);

// One dream_arr_annu call (src/dream_arr_annu.cc:514-634): Gr, focal and
// apod are host arrays, Ro, delay and h device pointers.
struct AnnuProblem {
  const double *Ro = nullptr;  int64_t No = 0;
  double dx = 0, dy = 0, dt = 0;  int64_t nt = 0;
  int delay_type = 0;  const double *delay = nullptr;
  double v = 1, cp = 1000;
  int64_t num_radii = 0;  const double *Gr = nullptr;
  int foc_met = 0;  const double *focal = nullptr;
  const double *apod = nullptr;  int do_apod = 0, apod_met = 0;  double apod_par = 0;
  int err_level = L_STOP;
  double alpha = 0;    // attenuation [dB/(cm MHz)]; > DBL_EPSILON selects it
  void *h = nullptr;   // nt x No, FP64 or FP32
};

// Annular array response: every disc through one SirPlan, then the ring
// superposition.  Synchronises `s`; *sir_err gets the reference's SIRError.
void run_arr_annu(const AnnuProblem &a, int out_dtype, hipStream_t s, int *sir_err,
                  int64_t *oob_count);

// Device parity helper: per-contribution sample indices of one
// (point, element) pair from the same device index routine.
void rsq_seed(const double *s, int64_t n, double *y);
void rect_indices(double xo, double yo, double zo, double a, double b, double dx, double dy,
                  double dt, double delay, double cp, int64_t *it_out, int64_t cap,
                  int64_t *count);

} // namespace dream
