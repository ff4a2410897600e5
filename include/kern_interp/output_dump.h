// Text dumps read by the reference's plotting scripts (no library counterpart
// there: the writers sit, commented out, in experiment_one.cpp:43-60,
// experiment_two.cpp:64-80 and experiment_three.cpp:211-226; the readers are
// output/domain_plot.py:79-80 and output/animation.py:47-62).
//   solution file: one line per domain point, "x,y,u" (solution_dimension 1)
//                  or "x,y,u,v" (2), in domain-point order;
//   boundary file: one line per boundary node, "x,y".
// Numbers are written as the reference's std::ostream << double does
// (6 significant digits, %g style).
#ifndef KERN_INTERP_OUTPUT_DUMP_H_
#define KERN_INTERP_OUTPUT_DUMP_H_

#include <string>
#include <vector>

#include "kern_interp/boundaries/boundary.h"
#include "kern_interp/ki_mat.h"

namespace kern_interp {

// false when the file cannot be written
bool write_solution_file(const std::string& path, const std::vector<double>& domain_points,
                         const ki_Mat& solution, int solution_dimension);
bool write_boundary_file(const std::string& path, const Boundary& boundary);
// one animation frame: <dir>/sol/<frame>.txt and <dir>/boundary/<frame>.txt
// (output/animation.py reads output/bake/...); the directories must exist
bool write_frame(const std::string& dir, int frame, const std::vector<double>& domain_points,
                 const ki_Mat& solution, int solution_dimension, const Boundary& boundary);

}  // namespace kern_interp

#endif  // KERN_INTERP_OUTPUT_DUMP_H_
