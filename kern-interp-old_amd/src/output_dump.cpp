// Text dumps for the plotting scripts (declared in kern_interp/output_dump.h).
#include "kern_interp/output_dump.h"

#include <fstream>

namespace kern_interp {

bool write_solution_file(const std::string& path, const std::vector<double>& pts,
                         const ki_Mat& sol, int sd) {
  std::ofstream out(path);
  if (!out) return false;
  const int n = pts.size() / 2;
  for (int p = 0; p < n; ++p) {
    out << pts[2 * p] << "," << pts[2 * p + 1];
    for (int c = 0; c < sd; ++c) out << "," << sol.get(sd * p + c, 0);
    out << "\n";
  }
  return (bool)out;
}

bool write_boundary_file(const std::string& path, const Boundary& b) {
  std::ofstream out(path);
  if (!out) return false;
  for (size_t i = 0; i + 1 < b.points.size(); i += 2) out << b.points[i] << "," << b.points[i + 1] << "\n";
  return (bool)out;
}

bool write_frame(const std::string& dir, int frame, const std::vector<double>& pts,
                 const ki_Mat& sol, int sd, const Boundary& b) {
  const std::string f = std::to_string(frame) + ".txt";
  return write_solution_file(dir + "/sol/" + f, pts, sol, sd) &&
         write_boundary_file(dir + "/boundary/" + f, b);
}

}  // namespace kern_interp
