// Houdini PGEOMETRY V5 ASCII writer: the text layout of
// util/houdini_geo/HoudiniFileDumpHelper.cpp:19-90 with the attributes of
// HoudiniFileDumpHelper.h:172-210 (v, color, mass), and the density colour
// map of houdini_file_saver.cpp:44-68.
//
// std::ostream prints a float with the default "%g"-style conversion at
// precision 6 of its double value; the per-point lines are produced with
// snprintf("%.6g") of the same doubles, which is the identical conversion,
// and are formatted in parallel chunks before being written in order.
#include "houdini_file_saver.h"

#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <thread>
#include <vector>

namespace {

struct point {
  float px, py, pz, vx, vy, vz, r, g, b, mass;
};

point make_point(const particle& p, float mass) {
  point q;
  q.px = p.position.s[0];
  q.py = p.position.s[1];
  q.pz = p.position.s[2];
  q.vx = p.velocity.s[0];
  q.vy = p.velocity.s[1];
  q.vz = p.velocity.s[2];
  float d = p.density;
  q.r = d > 1000.f && d <= 2000.f ? (d - 1000.f) / 1000.f : 0.f;
  q.g = d >= 0.f && d < 1000.f ? 1.f - d / 1000.f : 0.f;
  q.b = d >= 500.f && d <= 1000.f
            ? (d - 500.f) / 500.f
            : d >= 1000.f && d <= 1500.f ? 1.f - (d - 1000.f) / 500.f : 0.f;
  q.mass = mass;
  return q;
}

void append_g(std::string& out, float v) {
  char buf[48];
  int n = std::snprintf(buf, sizeof(buf), "%.6g", (double)v);
  out.append(buf, (size_t)n);
}

void format_points(const particle* p, size_t begin, size_t end, float mass, std::string& out) {
  out.reserve((end - begin) * 112);
  for (size_t i = begin; i < end; ++i) {
    point q = make_point(p[i], mass);
    append_g(out, q.px);
    out += ' ';
    append_g(out, q.py);
    out += ' ';
    append_g(out, q.pz);
    out += " 0 (";
    append_g(out, q.vx);
    out += ' ';
    append_g(out, q.vy);
    out += ' ';
    append_g(out, q.vz);
    out += '\t';
    append_g(out, q.r);
    out += ' ';
    append_g(out, q.g);
    out += ' ';
    append_g(out, q.b);
    out += '\t';
    append_g(out, q.mass);
    out += ")\n";
  }
}

}  // namespace

std::string clsph_zero_pad(int num) {
  std::ostringstream ss;
  ss.width(9);
  ss.fill('0');
  ss << num;
  std::string r = ss.str();
  if (r.length() > 7) r.erase(0, r.length() - 7);
  return r;
}

std::string clsph_houdini_geo_text(const particle* particles, unsigned int n, float mass) {
  std::string head;
  head += "PGEOMETRY V5\n";
  head += "NPoints " + std::to_string(n) + " NPrims 1\n";
  head += "NPointGroups 0 NPrimGroups 1\n";
  head += "NPointAttrib 3 NVertexAttrib 0 NPrimAttrib 2 NAttrib 0\n";
  head += "PointAttrib\n";
  head += "v 3 float 1 1 1\n";
  head += "color 3 float 1 1 1\n";
  head += "mass 1 float 1\n";

  unsigned hw = std::thread::hardware_concurrency();
  size_t chunks = n < 65536 ? 1 : (hw ? hw : 4);
  std::vector<std::string> parts(chunks);
  std::vector<std::thread> pool;
  for (size_t c = 0; c < chunks; ++c) {
    size_t b = (size_t)n * c / chunks, e = (size_t)n * (c + 1) / chunks;
    if (chunks == 1)
      format_points(particles, b, e, mass, parts[c]);
    else
      pool.emplace_back(format_points, particles, b, e, mass, std::ref(parts[c]));
  }
  for (auto& t : pool) t.join();

  std::string tail;
  tail += "PrimitiveAttrib\n";
  tail += "generator 1 index 1 location1\n";
  tail += "dopobject 1 index 1 /obj/AutoDopNetwork:1\n";
  tail += "Part " + std::to_string(n);
  for (unsigned int k = 0; k < n; ++k) {
    tail += ' ';
    tail += std::to_string(k);
  }
  tail += " [0\t0]\n";
  tail += "box_object1 unordered\n";
  tail += "1 1\n";
  tail += "beginExtra\n";
  tail += "endExtra\n";

  size_t total = head.size() + tail.size();
  for (auto& s : parts) total += s.size();
  std::string out;
  out.reserve(total);
  out += head;
  for (auto& s : parts) out += s;
  out += tail;
  return out;
}

int houdini_file_saver::writeFrameToFile(particle* particles,
                                         const simulation_parameters& parameters) {
  std::string name =
      frames_folder_prefix + "frames/frame" + clsph_zero_pad(++frame_count) + ".geo";
  std::ofstream f(name, std::ios::out | std::ios::trunc | std::ios::binary);
  if (f.is_open()) {
    std::string text =
        clsph_houdini_geo_text(particles, parameters.particles_count, parameters.particle_mass);
    f.write(text.data(), (std::streamsize)text.size());
    f.close();
  } else {
    std::cerr << "Error while writing to " << name << std::endl;
  }
  return 0;
}
