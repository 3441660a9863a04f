// Continuous-batching server behind fm_submit / fm_poll: requests (one
// periodic structure each) are queued as they arrive; each worker thread
// owns a context (HIP stream) with every kernel's model resident and
// repeatedly takes the oldest waiting requests whose device footprint fits
// the HBM budget left beside the batches in flight on all workers, builds
// them as one batch (one device batch per descriptor calculator), predicts
// and hands each request its own block of the results. A request's results
// do not depend on the other requests of its batch (per-structure reduction
// chunks, fixed summation orders), so they equal evaluating it alone.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fm_internal.hpp"

namespace {

struct Request {
  int64_t id = 0;
  double cell[9];
  std::vector<int32_t> species;
  std::vector<double> pos;
  int64_t need = 0;
  bool done = false;
  int rc = FM_OK;
  std::string error;
  std::vector<double> mean, var;  // var: n_kernels x N
};

struct Kernel {
  fm_b2_settings settings;
  std::vector<double> cutoffs;
  int kind = FM_DESCRIPTOR_B2;
  int nd = 0, nrad_h = 0, max_sparse = 0;
};

int check_rc(int rc) {
  if (rc != FM_OK) throw fm::Error(rc, fm_last_error());
  return rc;
}

}  // namespace

struct fm_server_s {
  std::vector<Kernel> kernels;
  double cutoff = 0, max_density = 0.1;
  int form = FM_FORM_CHOSEN, flags = FM_PREDICT_MEAN;
  int64_t budget = 0;
  int max_batch_atoms = 0;
  std::vector<fm_context> ctxs;
  std::vector<std::vector<fm_model>> models;  // [worker][kernel]
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> queue;
  std::unordered_map<int64_t, std::unique_ptr<Request>> requests;
  int64_t next_id = 1, inflight = 0;
  bool stop = false;
  std::vector<std::thread> workers;

  // Device bytes of a batch: neighbour list, bond list and bond forces per
  // neighbour; descriptors, gradient rows, density coefficients and their
  // adjoint, int8 slices, w and k per atom, for every kernel.
  int64_t bytes(int64_t n_atoms, int64_t n_neighbors) const {
    int64_t esz = form == FM_FORM_FP64 ? 8 : 4, total = 0;
    for (const Kernel &k : kernels)
      total += n_neighbors * (56 + 3 * esz) +
               n_atoms * (2 * esz * k.nd + 2 * esz * k.nrad_h + 5 * k.nd + 12 * k.max_sparse +
                          256);
    return total;
  }

  int64_t estimate(const double *c, int n) const {
    double vol = std::fabs(c[0] * (c[4] * c[8] - c[5] * c[7]) - c[1] * (c[3] * c[8] - c[5] * c[6]) +
                           c[2] * (c[3] * c[7] - c[4] * c[6]));
    double rho = std::max(max_density, vol > 0 ? n / vol : 0.0);
    double nb = rho * 4.0 / 3.0 * M_PI * cutoff * cutoff * cutoff;
    return bytes(n, (int64_t)std::ceil(n * nb));
  }

  // The oldest waiting requests that fit beside the batches in flight.
  std::vector<Request *> admit(int64_t &used) {
    std::vector<Request *> batch;
    int64_t atoms = 0;
    used = 0;
    while (!queue.empty()) {
      Request *r = queue.front();
      int64_t n = (int64_t)r->species.size();
      if (inflight + used + r->need > budget) break;
      if (!batch.empty() && max_batch_atoms > 0 && atoms + n > max_batch_atoms) break;
      batch.push_back(r);
      queue.pop_front();
      used += r->need;
      atoms += n;
    }
    return batch;
  }

  void run_batch(int w, std::vector<Request *> &reqs, int64_t &used) {
    int ns = (int)reqs.size();
    std::vector<double> cells(9 * (size_t)ns), pos;
    std::vector<int32_t> n_atoms(ns), species;
    for (int i = 0; i < ns; i++) {
      std::memcpy(&cells[9 * (size_t)i], reqs[i]->cell, 9 * sizeof(double));
      n_atoms[i] = (int32_t)reqs[i]->species.size();
      species.insert(species.end(), reqs[i]->species.begin(), reqs[i]->species.end());
      pos.insert(pos.end(), reqs[i]->pos.begin(), reqs[i]->pos.end());
    }
    int64_t n = (int64_t)species.size(), n_out = 0;
    for (int i = 0; i < ns; i++) n_out += 1 + 3 * (int64_t)n_atoms[i] + 6;
    std::vector<fm_structure> hs(kernels.size(), nullptr);
    struct Release {
      std::vector<fm_structure> &h;
      ~Release() {
        for (fm_structure s : h)
          if (s) fm_structure_destroy(s);
      }
    } release{hs};
    for (size_t k = 0; k < kernels.size(); k++) {
      check_rc(fm_batch_create(ctxs[w], ns, cells.data(), n_atoms.data(), species.data(),
                               pos.data(), cutoff, &hs[k]));
      if (kernels[k].kind == FM_DESCRIPTOR_B3)
        check_rc(fm_structure_compute_b3(hs[k], &kernels[k].settings, form));
      else
        check_rc(fm_structure_compute_b2(hs[k], &kernels[k].settings, form));
    }
    // The measured neighbour count replaces the estimate.
    int sweep = 0;
    int64_t nnb = 0;
    double vol = 0, ssc = 0;
    check_rc(fm_structure_info(hs[0], &sweep, &nnb, &vol, &ssc));
    int64_t real = bytes(n, nnb);
    {
      std::lock_guard<std::mutex> g(mu);
      inflight += real - used;
    }
    used = real;
    bool local = (flags & FM_PREDICT_LOCAL) != 0;
    std::vector<double> mean(n_out, 0.0), out(n_out), var(local ? (size_t)n * kernels.size() : 0);
    for (size_t k = 0; k < kernels.size(); k++) {
      check_rc(fm_predict(models[w][k], hs[k], flags, form, out.data(),
                          local ? var.data() + k * (size_t)n : nullptr, nullptr));
      for (int64_t q = 0; q < n_out; q++) mean[q] += out[q];
    }
    int64_t mo = 0, ao = 0;
    for (int i = 0; i < ns; i++) {
      int na = n_atoms[i];
      Request *r = reqs[i];
      r->mean.assign(mean.begin() + mo, mean.begin() + mo + 1 + 3 * (int64_t)na + 6);
      if (local) {
        r->var.resize((size_t)na * kernels.size());
        for (size_t k = 0; k < kernels.size(); k++)
          std::memcpy(&r->var[k * na], &var[k * (size_t)n + ao], na * sizeof(double));
      }
      mo += 1 + 3 * (int64_t)na + 6;
      ao += na;
    }
  }

  void work(int w) {
    (void)hipSetDevice(ctxs[w]->device);
    for (;;) {
      std::vector<Request *> reqs;
      int64_t used = 0;
      {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
          if (stop && queue.empty()) return;
          reqs = admit(used);
          if (!reqs.empty()) break;
          cv_work.wait(lk);
        }
        inflight += used;
      }
      int rc = FM_OK;
      std::string err;
      try {
        run_batch(w, reqs, used);
      } catch (const fm::Error &e) {
        rc = e.code;
        err = e.what();
      } catch (const std::exception &e) {
        rc = FM_ERR_INVALID;
        err = e.what();
      }
      {
        std::lock_guard<std::mutex> g(mu);
        for (Request *r : reqs) {
          r->done = true;
          r->rc = rc;
          r->error = err;
        }
        inflight -= used;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }
};

extern "C" {

int fm_server_create(int device, const fm_server_desc *d, fm_server *out) {
  try {
    FM_CHECK(d && out, FM_ERR_INVALID, "null argument");
    FM_CHECK(d->n_kernels >= 1 && d->models && d->descriptors, FM_ERR_INVALID,
             "a server needs at least one kernel with its descriptor settings");
    FM_CHECK(d->n_workers >= 1 && d->n_workers <= 16, FM_ERR_INVALID,
             "workers must be between 1 and 16");
    FM_CHECK(d->form == FM_FORM_FP64 || d->form == FM_FORM_CHOSEN, FM_ERR_INVALID, "bad form");
    FM_CHECK(d->cutoff > 0 && d->hbm_budget > 0, FM_ERR_INVALID, "bad cutoff or budget");
    auto srv = std::make_unique<fm_server_s>();
    srv->cutoff = d->cutoff;
    srv->max_density = d->max_density > 0 ? d->max_density : 0.1;
    srv->form = d->form;
    srv->flags = d->flags & (FM_PREDICT_MEAN | FM_PREDICT_LOCAL);
    srv->budget = d->hbm_budget;
    srv->max_batch_atoms = d->max_batch_atoms;
    for (int k = 0; k < d->n_kernels; k++) {
      Kernel kk;
      kk.settings = d->descriptors[k];
      int nos = kk.settings.n_species;
      FM_CHECK(nos >= 1 && kk.settings.cutoffs, FM_ERR_INVALID, "bad descriptor settings");
      kk.cutoffs.assign(kk.settings.cutoffs, kk.settings.cutoffs + nos * nos);
      kk.settings.cutoffs = kk.cutoffs.data();
      kk.kind = d->descriptor_kinds ? d->descriptor_kinds[k] : FM_DESCRIPTOR_B2;
      FM_CHECK(kk.kind == FM_DESCRIPTOR_B2 || kk.kind == FM_DESCRIPTOR_B3, FM_ERR_INVALID,
               "descriptor kind must be FM_DESCRIPTOR_B2 or FM_DESCRIPTOR_B3");
      int nr = nos * kk.settings.n_max, L = kk.settings.l_max;
      kk.nd = d->models[k].n_descriptors;
      kk.nrad_h = nr * (L + 1) * (L + 1);
      for (int t = 0; t < d->models[k].n_species; t++)
        kk.max_sparse = std::max(kk.max_sparse, (int)d->models[k].n_sparse_by_type[t]);
      srv->kernels.push_back(std::move(kk));
    }
    // Each worker: its own context and every kernel's model on it. On a
    // failure here, what was made is released.
    struct Undo {
      fm_server_s *s;
      bool armed = true;
      ~Undo() {
        if (!armed) return;
        for (auto &ms : s->models)
          for (fm_model m : ms) fm_model_destroy(m);
        for (fm_context c : s->ctxs) fm_context_destroy(c);
      }
    } undo{srv.get()};
    for (int w = 0; w < d->n_workers; w++) {
      fm_context c = nullptr;
      check_rc(fm_context_create(device, &c));
      srv->ctxs.push_back(c);
      srv->models.emplace_back();
      for (int k = 0; k < d->n_kernels; k++) {
        fm_model m = nullptr;
        check_rc(fm_model_create(c, &d->models[k], &m));
        srv->models.back().push_back(m);
      }
    }
    undo.armed = false;
    fm_server_s *s = srv.release();
    for (int w = 0; w < d->n_workers; w++) s->workers.emplace_back([s, w] { s->work(w); });
    *out = s;
    return FM_OK;
  } catch (const fm::Error &e) {
    fm::set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    fm::set_last_error(e.what());
    return FM_ERR_INVALID;
  }
}

int fm_server_destroy(fm_server s) {
  if (!s) return FM_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  for (auto &ms : s->models)
    for (fm_model m : ms) fm_model_destroy(m);
  for (fm_context c : s->ctxs) fm_context_destroy(c);
  delete s;
  return FM_OK;
}

int fm_submit(fm_server s, const double *cell, int n_atoms, const int32_t *species,
              const double *positions, int64_t *request) {
  try {
    FM_CHECK(s && cell && request && n_atoms >= 0, FM_ERR_INVALID, "null argument");
    FM_CHECK(n_atoms == 0 || (species && positions), FM_ERR_INVALID,
             "null species or positions");
    auto r = std::make_unique<Request>();
    std::memcpy(r->cell, cell, sizeof(r->cell));
    r->species.assign(species, species + n_atoms);
    r->pos.assign(positions, positions + 3 * (size_t)n_atoms);
    r->need = s->estimate(cell, n_atoms);
    FM_CHECK(r->need <= s->budget, FM_ERR_CAPACITY,
             "request of " + std::to_string(n_atoms) + " atoms exceeds the HBM budget");
    {
      std::lock_guard<std::mutex> g(s->mu);
      FM_CHECK(!s->stop, FM_ERR_STATE, "server is shutting down");
      r->id = s->next_id++;
      *request = r->id;
      s->queue.push_back(r.get());
      s->requests.emplace(r->id, std::move(r));
    }
    s->cv_work.notify_one();
    return FM_OK;
  } catch (const fm::Error &e) {
    fm::set_last_error(e.what());
    return e.code;
  }
}

int fm_poll(fm_server s, int64_t request, int timeout_ms, int *done, double *mean_efs,
            double *local_var) {
  try {
    FM_CHECK(s && done, FM_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(request);
    FM_CHECK(it != s->requests.end(), FM_ERR_INVALID, "unknown request");
    Request *r = it->second.get();
    if (timeout_ms < 0)
      s->cv_done.wait(lk, [r] { return r->done; });
    else if (timeout_ms > 0)
      s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [r] { return r->done; });
    *done = r->done ? 1 : 0;
    if (!r->done) return FM_OK;
    std::unique_ptr<Request> owned = std::move(it->second);
    s->requests.erase(it);
    lk.unlock();
    FM_CHECK(owned->rc == FM_OK, owned->rc, owned->error);
    if (mean_efs) std::memcpy(mean_efs, owned->mean.data(), owned->mean.size() * sizeof(double));
    if (local_var && !owned->var.empty())
      std::memcpy(local_var, owned->var.data(), owned->var.size() * sizeof(double));
    return FM_OK;
  } catch (const fm::Error &e) {
    fm::set_last_error(e.what());
    return e.code;
  }
}

}  // extern "C"
