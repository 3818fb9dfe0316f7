// pn_track.hip -- island tracking over time: the bookkeeping of Full_Analysis
// (islands.cpp:788-835), Old_islands (:651-705) and print_bubble_history
// (:707-736) on top of the GPU island catalog (pn_islands).
//
// Each update is one Full_Analysis call.  The tracked islands are sorted by
// their latest volume, largest first, with std::sort and the reference's
// comparison (islands.cpp:29-32), so islands of equal volume come out in the
// same order as in the reference.  An island whose last peak site left the
// mask is retired (written to non_critical_islands.txt and islands.txt).  The
// survivors, in sorted order, re-grow from their last peak: the first to reach
// a component claims it and records a new history entry, a later one whose
// peak lies in an already claimed component has merged and is dropped.  The
// components nobody claimed are new islands, numbered in raster order of
// their first site.  The reference re-grows by a BFS from the old peak; here
// the component containing that site comes from the GPU labelling
// (pn_islands_lookup) and its statistics from the catalog: the same partition
// and the same sums up to summation order.  A unique maximum is the peak
// whatever the seed; when the component holds equal maxima the peak is the
// first of them in the BFS order from the old peak (pn_islands_seeded_peak),
// as the reference's re-growth finds it.
//
// Histories keep at most five entries: the first two and the latest three
// (islands.cpp:255-267).  They live in a pool of five-entry slots, so the
// list itself holds small records that the sort and the rebuild of the list
// move cheaply (a 256^3 lattice at the reference's threshold tracks some
// 400,000 islands).  The catalog is copied into page-locked memory.  Text
// outputs use the reference's stream formats.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <new>
#include <string>
#include <vector>

#include "pn_internal.h"

struct pn_tracker {
  // a tracked island; its history is slot `slot` of the pool
  struct Isl {
    long long index = 0;
    int disp = 1;
    int status = 0;
    int nh = 0;
    int slot = 0;
  };
  std::vector<Isl> found;
  std::vector<pn_track_entry> pool;  // PN_TRACK_HIST entries per slot
  std::vector<int> free_slots;
  long long number_of_islands = 0;
  // the last update's components (raster numbering), page-locked
  pn_island *cat = nullptr;
  int cat_cap = 0;
  int cat_n = 0;
  int cap_hint = 1024;  // catalog room for the next update: the last count plus a quarter
  std::string err;

  ~pn_tracker() {
    if (cat) hipHostFree(cat);
  }
  pn_track_entry *hist(const Isl &s) { return pool.data() + (size_t)s.slot * PN_TRACK_HIST; }
  const pn_track_entry *hist(const Isl &s) const { return pool.data() + (size_t)s.slot * PN_TRACK_HIST; }
  const pn_track_entry &back(const Isl &s) const { return hist(s)[s.nh - 1]; }
  const pn_track_entry &front(const Isl &s) const { return hist(s)[0]; }
  // islands.cpp:255-267: the first two entries and the latest three
  void push(Isl &s, const pn_track_entry &e) {
    pn_track_entry *h = hist(s);
    if (s.nh == PN_TRACK_HIST) {
      for (int j = 2; j + 1 < PN_TRACK_HIST; j++) h[j] = h[j + 1];
      h[PN_TRACK_HIST - 1] = e;
    } else {
      h[s.nh++] = e;
    }
  }
  int new_slot() {
    if (!free_slots.empty()) {
      const int k = free_slots.back();
      free_slots.pop_back();
      return k;
    }
    pool.resize(pool.size() + PN_TRACK_HIST);
    return (int)(pool.size() / PN_TRACK_HIST) - 1;
  }
};

static int tfail(pn_tracker *t, int code, const std::string &m) {
  if (t) t->err = m;
  pn_set_thread_error(m.c_str());
  return code;
}

extern "C" int pn_tracker_create(pn_tracker **out) {
  if (!out) return tfail(nullptr, PN_ERR_ARG, "pn_tracker_create: NULL argument");
  *out = new (std::nothrow) pn_tracker();
  return *out ? PN_OK : tfail(nullptr, PN_ERR_NOMEM, "pn_tracker_create: out of host memory");
}

extern "C" int pn_tracker_destroy(pn_tracker *t) {
  delete t;
  return PN_OK;
}

extern "C" const char *pn_tracker_last_error(const pn_tracker *t) {
  return t ? t->err.c_str() : pn_last_error(nullptr);
}

extern "C" int pn_tracker_catalog(const pn_tracker *t, pn_catalog *out) {
  if (!t || !out || (out->capacity > 0 && !out->islands))
    return tfail(const_cast<pn_tracker *>(t), PN_ERR_ARG, "pn_tracker_catalog: bad argument");
  out->count = t->cat_n;
  const int n = std::min(out->count, out->capacity);
  if (n > 0) std::copy(t->cat, t->cat + n, out->islands);
  if (out->count > out->capacity)
    return tfail(const_cast<pn_tracker *>(t), PN_ERR_CAPACITY, "pn_tracker_catalog: capacity below the catalog");
  return PN_OK;
}

extern "C" int pn_tracker_count(const pn_tracker *t, int *n) {
  if (!t || !n) return tfail(const_cast<pn_tracker *>(t), PN_ERR_ARG, "pn_tracker_count: NULL argument");
  *n = (int)t->found.size();
  return PN_OK;
}

extern "C" int pn_tracker_get(const pn_tracker *t, int i, pn_tracked_island *out) {
  if (!t || !out || i < 0 || i >= (int)t->found.size())
    return tfail(const_cast<pn_tracker *>(t), PN_ERR_ARG, "pn_tracker_get: bad argument");
  const pn_tracker::Isl &s = t->found[i];
  out->index = s.index;
  out->status = s.status;
  out->disp = s.disp;
  out->nhist = s.nh;
  const pn_track_entry *h = t->hist(s);
  for (int j = 0; j < out->nhist; j++) out->hist[j] = h[j];
  return PN_OK;
}

namespace {

// point::x/y/z (fieldanalysis.cpp:19-32) for one process (id 0)
struct Coords {
  long long N;
  long long x(long long pos) const { return (pos % (N * N)) % N; }
  long long y(long long pos) const { return (pos % (N * N)) / N; }
  long long z(long long pos) const { return -1 + pos / (N * N); }
};

// The reference writes these rows with ostream << (default precision 6,
// that is printf's %g, and plain integers) ending in std::endl.  Here each
// row is formatted with %g into a buffer and written once, ending in '\n':
// the same bytes, without the per-field stream formatting and the flush (a
// write system call) per row.
void write_entry(std::ofstream &o, const pn_track_entry &e, long long index, const Coords &c) {
  char b[512];
  const int n = std::snprintf(b, sizeof b, "%g %lld %lld %lld        %g %g %g %g %g %g %lld\n", e.at_time,
                              c.x(e.center_pos), c.y(e.center_pos), c.z(e.center_pos), e.peak_value,
                              e.avg_phi, e.volume, e.grad_en, e.pot_en, e.kin_en, index);
  o.write(b, n);
}

// the islands.txt row (islands.cpp:683,717)
void write_init_row(std::ofstream &o, const pn_track_entry &e0, long long index, int status, double thres) {
  const double r = std::pow(e0.volume * 3. / 4. / std::acos(-1.), 1. / 3.);
  char b[256];
  const int n = std::snprintf(b, sizeof b, "%g %g %g %lld %d %g\n", r, e0.peak_value, e0.at_time, index,
                              status, r / std::sqrt(std::log(e0.peak_value * e0.peak_value / thres / thres)));
  o.write(b, n);
}

// a tracked island's sort key: its latest volume, and its place in the list
struct VolKey {
  double v;
  int i;
};

}  // namespace

static pn_track_entry entry_of(const pn_island &I, double tau) {
  pn_track_entry e;
  e.at_time = tau;
  e.center_pos = I.peak_pos;
  e.peak_value = I.peak_value;
  e.avg_phi = I.avg_phi;
  e.volume = I.volume;
  e.pot_en = I.pot_en;
  e.kin_en = I.kin_en;
  e.grad_en = I.grad_en;
  return e;
}

// the catalog array (page-locked) with room for at least `n` rows
static int catalog_room(pn_tracker *t, int n) {
  if (n <= t->cat_cap) return PN_OK;
  if (t->cat) hipHostFree(t->cat);
  t->cat = nullptr;
  t->cat_cap = 0;
  t->cat_n = 0;
  void *p = nullptr;
  const int rc = pn_host_alloc(sizeof(pn_island) * (size_t)n, &p);
  if (rc != PN_OK) return rc;
  t->cat = (pn_island *)p;
  t->cat_cap = n;
  return PN_OK;
}

static int tracker_update(pn_tracker *t, pn_ctx *c, int mode, double thr, double tau,
                          const char *outdir) {
  // the current components (room for the last count plus a quarter, so that
  // the analysis normally runs once)
  int rc = catalog_room(t, t->cap_hint);
  if (rc != PN_OK) return tfail(t, rc, std::string("pn_tracker_update: ") + pn_last_error(c));
  pn_catalog pc{t->cat_cap, 0, 0.0, t->cat, nullptr};
  rc = pn_islands(c, mode, thr, &pc);
  if (rc == PN_ERR_CAPACITY) {
    const int need = pc.count + pc.count / 4;
    if ((rc = catalog_room(t, need)) != PN_OK)
      return tfail(t, rc, std::string("pn_tracker_update: ") + pn_last_error(c));
    pc = pn_catalog{t->cat_cap, 0, 0.0, t->cat, nullptr};
    rc = pn_islands(c, mode, thr, &pc);
  }
  if (rc != PN_OK) {
    t->cat_n = 0;
    return tfail(t, rc, std::string("pn_tracker_update: ") + pn_last_error(c));
  }
  t->cat_n = pc.count;
  t->cap_hint = std::max(1024, pc.count + pc.count / 4);
  const pn_island *cat = t->cat;
  const int ncat = pc.count;

  // largest latest volume first (islands.cpp:29-32,797): std::sort with the
  // reference's comparison, so ties keep its order.  The keys are sorted
  // with the list positions beside them: std::sort makes the same
  // comparisons, with the same outcomes, as on the islands themselves, so
  // it permutes them the same way.
  const int n = (int)t->found.size();
  std::vector<VolKey> ord(n);
  for (int i = 0; i < n; i++) ord[i] = VolKey{t->back(t->found[i]).volume, i};
  std::sort(ord.begin(), ord.end(), [](const VolKey &a, const VolKey &b) { return a.v > b.v; });

  // component of each tracked island's last peak (-1: the site left the mask)
  std::vector<long long> peaks(n);
  std::vector<int> comp(n, -1);
  for (int i = 0; i < n; i++) peaks[i] = t->back(t->found[ord[i].i]).center_pos;
  if (n > 0 && (rc = pn_islands_lookup(c, n, peaks.data(), comp.data())) != PN_OK)
    return tfail(t, rc, std::string("pn_tracker_update: ") + pn_last_error(c));

  const Coords co{c->p.N};
  std::ofstream safe, init, bubbles;
  const bool files = outdir != nullptr;
  if (files) {
    // Old_islands opens (and so creates) these two on every call (:663-664)
    safe.open(std::string(outdir) + "/non_critical_islands.txt", std::ios::app);
    init.open(std::string(outdir) + "/islands.txt", std::ios::app);
    if (!safe || !init) return tfail(t, PN_ERR_ARG, std::string("pn_tracker_update: cannot write to ") + outdir);
  }

  // Old_islands: retire the islands whose peak site is no longer in the mask
  if (files)
    for (int i = 0; i < n; i++) {
      if (comp[i] >= 0) continue;
      const pn_tracker::Isl &s = t->found[ord[i].i];
      const pn_track_entry *h = t->hist(s);
      if (s.disp) {
        for (int j = 0; j < s.nh; j++) write_entry(safe, h[j], s.index, co);
        safe << '\n';
      }
      write_init_row(init, h[0], s.index, s.status, thr);
    }

  // survivors re-grow from their last peak in sorted order; a claimed
  // component means a merge.  With equal maxima in the component the peak is
  // the first of them in breadth-first order from that old peak.
  std::vector<char> claimed((size_t)ncat, 0);
  std::vector<pn_tracker::Isl> next;
  next.reserve((size_t)n + ncat);
  for (int i = 0; i < n; i++) {
    const int k = comp[i];
    pn_tracker::Isl s = t->found[ord[i].i];
    if (k < 0 || claimed[k]) {  // retired, or merged into an island earlier in the order
      t->free_slots.push_back(s.slot);
      continue;
    }
    claimed[k] = 1;
    pn_track_entry e = entry_of(cat[k], tau);
    if (cat[k].peak_ties > 0 &&
        (rc = pn_islands_seeded_peak(c, k, t->back(s).center_pos, &e.center_pos, &e.peak_value)) != PN_OK)
      return tfail(t, rc, std::string("pn_tracker_update: ") + pn_last_error(c));
    t->push(s, e);
    s.status = 0;
    if (e.grad_en + e.pot_en < 0.) {  // is_bubble (islands.cpp:61-65)
      if (files) {
        // print_bubble_history (:707-736) opens expanding_bubbles.txt when called
        if (!bubbles.is_open()) bubbles.open(std::string(outdir) + "/expanding_bubbles.txt", std::ios::app);
        const pn_track_entry *h = t->hist(s);
        if (s.disp) write_init_row(init, h[0], s.index, 1, thr);
        for (int j = 1; j < s.nh; j++) write_entry(bubbles, h[j], s.index, co);
        bubbles << '\n';
      }
      s.status = 1;
    }
    if (s.status == 0 && e.at_time - t->front(s).at_time > 40.)  // is_oscillon (:67-70)
      s.status = 2;
    next.push_back(s);
  }

  // new islands: unclaimed components in raster order of their first site
  // (their peaks, ties included, are the catalog's: BFS from that site)
  for (int k = 0; k < ncat; k++) {
    if (claimed[k]) continue;
    pn_tracker::Isl s;
    s.index = t->number_of_islands++;
    s.disp = 1;
    s.status = 0;
    s.slot = t->new_slot();
    t->push(s, entry_of(cat[k], tau));
    next.push_back(s);
  }
  t->found.swap(next);
  return PN_OK;
}

extern "C" int pn_tracker_update(pn_tracker *t, pn_ctx *c, int mode, double thr, double tau,
                                 const char *outdir) {
  if (!t || !c) return tfail(t, PN_ERR_ARG, "pn_tracker_update: NULL argument");
  try {
    return tracker_update(t, c, mode, thr, tau, outdir);
  } catch (const std::bad_alloc &) {
    return tfail(t, PN_ERR_NOMEM, "pn_tracker_update: out of host memory");
  }
}
