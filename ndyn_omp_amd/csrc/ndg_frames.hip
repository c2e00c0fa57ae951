// ndg_frames.hip -- trajectory frames in the reference's text formats, formatted on the host.
//
// write_model (src/read_input.f90:394-431) writes one frame per ConfOutFreq steps with
// Fortran edit descriptors.  The formatter below reproduces those descriptors byte for byte
// (Fw.d and ENw.d round the exact binary value to nearest, ties to even, as the Fortran
// runtime does; fields that do not fit are filled with '*'), splits the atoms over host
// threads and writes the frame with one fwrite.  A writer thread per handle formats queued
// frames while the caller keeps stepping the device.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ndg_frames.h"

namespace ndg {

namespace {

// right-justify s (n chars) in a field of w; '*' * w if it does not fit
inline void field(std::string &o, const char *s, int n, int w) {
  if (n > w) {
    o.append((size_t)w, '*');
    return;
  }
  o.append((size_t)(w - n), ' ');
  o.append(s, (size_t)n);
}

// Iw
inline void put_i(std::string &o, long long v, int w) {
  char b[32];
  const int n = snprintf(b, sizeof b, "%lld", v);
  field(o, b, n, w);
}

// Fw.d of a double (correctly rounded by the C library: nearest, ties to even)
inline void put_f(std::string &o, double x, int w, int d) {
  char b[400];
  if (!isfinite(x)) {
    const char *t = isnan(x) ? "NaN" : (x < 0 ? "-Inf" : "Inf");
    field(o, t, (int)strlen(t), w);
    return;
  }
  const int n = snprintf(b, sizeof b, "%.*f", d, x);
  field(o, b, n, w);
}

// F9.5 of a single-precision value: x * 1e5 is exact in double (24 + 17 bits), so rint
// (nearest, ties to even) rounds the exact value
inline void put_f95(std::string &o, float xf) {
  const double x = (double)xf;
  if (!(fabs(x) < 1.0e4)) {
    put_f(o, x, 9, 5);
    return;
  }
  const double r = rint(fabs(x) * 1.0e5);
  long long m = (long long)r;
  char b[24];
  int n = 0;
  if (signbit(x)) b[n++] = '-';
  const long long ip = m / 100000, fp = m % 100000;
  n += snprintf(b + n, sizeof b - n, "%lld.%05lld", ip, fp);
  field(o, b, n, 9);
}

// ENw.d: engineering notation, exponent a multiple of 3, 1 to 3 digits before the point.
// The digits come from the C library's correctly rounded %e at d + (integer digits)
// significant digits.  Exponents beyond 99 drop the 'E' (Fortran's default exponent form).
inline void put_en(std::string &o, double x, int w, int d) {
  char b[64], t[64];
  if (!isfinite(x)) {
    const char *s = isnan(x) ? "NaN" : (x < 0 ? "-Infinity" : "Infinity");
    field(o, s, (int)strlen(s), w);
    return;
  }
  const bool neg = signbit(x);
  const double a = fabs(x);
  int e3 = 0;
  std::string ip = "0", fr(d, '0');
  if (a != 0.0) {
    snprintf(t, sizeof t, "%.17e", a);
    const int et = atoi(strchr(t, 'e') + 1);          // exponent of the exact value
    e3 = (int)floor(et / 3.0) * 3;
    const int sig = (et - e3 + 1) + d;
    snprintf(t, sizeof t, "%.*e", sig - 1, a);
    const int e1 = atoi(strchr(t, 'e') + 1);
    if (e1 == et) {
      std::string dg;
      for (const char *c = t; *c && *c != 'e'; ++c)
        if (*c >= '0' && *c <= '9') dg.push_back(*c);
      const int ni = et - e3 + 1;
      ip = dg.substr(0, ni);
      fr = dg.substr(ni, d);
    } else {
      // rounded up to the next power of ten, 10^e1
      if (e1 - e3 >= 3) {
        e3 += 3;
        ip = "1";
      } else {
        ip = "1" + std::string(e1 - e3, '0');
      }
    }
  }
  int n = 0;
  if (neg) b[n++] = '-';
  n += snprintf(b + n, sizeof b - n, "%s.%s", ip.c_str(), fr.c_str());
  const int ae = e3 < 0 ? -e3 : e3;
  if (ae <= 99)
    n += snprintf(b + n, sizeof b - n, "E%c%02d", e3 < 0 ? '-' : '+', ae);
  else
    n += snprintf(b + n, sizeof b - n, "%c%03d", e3 < 0 ? '-' : '+', ae);
  field(o, b, n, w);
}

// Aw of a blank-padded name of length len
inline void put_a(std::string &o, const char *s, int len, int w) {
  if (len >= w) {
    o.append(s, (size_t)w);
  } else {
    o.append((size_t)(w - len), ' ');
    o.append(s, (size_t)len);
  }
}

void format_atoms(std::string &o, const FrameData &f, int i0, int i1) {
  const int np = f.np;
  if (f.format == NDG_FRAME_DUMP) {
    o.reserve((size_t)(i1 - i0) * 38);
    for (int i = i0; i < i1; ++i) {
      put_i(o, i + 1, 7);          // (i7,x,i2,3f9.5)
      o.push_back(' ');
      put_i(o, f.typ[i], 2);
      for (int a = 0; a < 3; ++a) put_f95(o, (float)f.pos[3 * i + a] + 0.0f);
      o.push_back('\n');
    }
  } else {
    o.reserve((size_t)(i1 - i0) * 104);
    for (int i = i0; i < i1; ++i) {
      const int t = f.typ[i];        // (A4,x,7EN14.4)
      put_a(o, f.names + (size_t)(t - 1) * f.name_len, f.name_len, 4);
      o.push_back(' ');
      for (int a = 0; a < 3; ++a) put_en(o, f.pos[3 * i + a], 14, 4);
      for (int a = 0; a < 3; ++a) put_en(o, f.vel[3 * i + a], 14, 4);
      put_en(o, f.pene[i], 14, 4);
      o.push_back('\n');
    }
  }
  (void)np;
}

}  // namespace

void format_frame(const FrameData &f, std::string &out, int nthreads) {
  out.clear();
  if (f.format == NDG_FRAME_DUMP) {
    out += "ITEM: TIMESTEP\n";
    put_i(out, f.step, 7);
    out += "\nITEM: NUMBER OF ATOMS\n";
    put_i(out, f.np, 7);
    out += "\nITEM: BOX BOUNDS pp pp pp\n";
    for (int a = 0; a < 3; ++a) {
      put_f(out, 0.0, 12, 4);
      put_f(out, f.box[a], 12, 4);
      out.push_back('\n');
    }
    out += "ITEM: ATOMS id type xs ys zs\n";
  } else {
    put_i(out, f.np, 8);
    out += "\n\n";
  }
  const int per = 4096;
  int nt = std::max(1, std::min(nthreads, (f.np + per - 1) / per));
  if (nt == 1) {
    std::string body;
    format_atoms(body, f, 0, f.np);
    out += body;
    return;
  }
  std::vector<std::string> parts(nt);
  std::vector<std::thread> th;
  for (int k = 0; k < nt; ++k) {
    const int i0 = (int)((long long)f.np * k / nt), i1 = (int)((long long)f.np * (k + 1) / nt);
    th.emplace_back([&, k, i0, i1] { format_atoms(parts[k], f, i0, i1); });
  }
  for (auto &t : th) t.join();
  size_t tot = out.size();
  for (auto &p : parts) tot += p.size();
  out.reserve(tot);
  for (auto &p : parts) out += p;
}

int format_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(h ? h : 1u, 16u));
}

}  // namespace ndg

extern "C" int ndg_format_frame(int format, int np, const double *pos, const double *vel,
                                const double *pene, const int *typ, const char *names,
                                int name_len, const double box[3], int step, char *buf,
                                size_t cap, size_t *len) {
  if ((format != NDG_FRAME_XYZ && format != NDG_FRAME_DUMP) || np < 0 || !pos || !typ || !len ||
      (format == NDG_FRAME_XYZ && (!vel || !pene || !names || name_len < 1)) ||
      (format == NDG_FRAME_DUMP && !box))
    return NDG_ERR_ARG;
  ndg::FrameData f;
  f.format = format;
  f.np = np;
  f.pos = pos;
  f.vel = vel;
  f.pene = pene;
  f.typ = typ;
  f.names = names;
  f.name_len = name_len;
  if (box)
    for (int a = 0; a < 3; ++a) f.box[a] = box[a];
  f.step = step;
  std::string s;
  ndg::format_frame(f, s, ndg::format_threads());
  *len = s.size();
  if (!buf) return NDG_OK;
  if (cap < s.size()) return NDG_ERR_ARG;
  memcpy(buf, s.data(), s.size());
  return NDG_OK;
}
