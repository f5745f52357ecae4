// ba_plugin.cc -- the map-optimization console plugin over the visual BA
// library (include/maplab_ba_plugin.h). Commands and flag names follow
// console-plugins/map-optimization-plugin/src/optimizer-plugin.cc:27-55 and
// the flags of solver-options.cc:3-52, outlier-rejection-solver.cc:9-21,
// vi-optimization-builder.cc:10-97 and landmark-quality-metrics.cc:11-33.
#include "../../include/maplab_ba_plugin.h"

#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/maplab_ba_optimizer.h"

namespace {

struct Command {
  std::vector<std::string> names;
  std::function<int()> callback;
  std::string help;
};

class OptimizerPlugin {
 public:
  explicit OptimizerPlugin(mlba_console* console) : console_(console) {
    addCommand({"optimize_visual", "optv"}, [this] { return optimize(/*inertial=*/false); },
               "Visual optimization over the selected missions (per default all).");
    addCommand({"optimize_visual_inertial", "optvi"}, [this] { return optimize(/*inertial=*/true); },
               "Visual-inertial optimization over the selected missions (per default all).");
    name_list_ = names();
  }

  const char* id() const { return "map-optimization"; }
  std::vector<std::string> names() const {
    std::vector<std::string> out;
    for (const Command& c : commands_)
      for (const std::string& n : c.names) out.push_back(n);
    return out;
  }
  const std::string& message() const { return message_; }

  int process(const std::string& line) {
    message_.clear();
    std::istringstream in(line);
    std::vector<std::string> tokens;
    for (std::string t; in >> t;) tokens.push_back(t);
    if (tokens.empty()) return fail(MLBA_CMD_STUPID_USER_ERROR, "empty command");
    flags_.clear();
    for (size_t i = 1; i < tokens.size(); ++i) {
      const std::string& t = tokens[i];
      if (t.rfind("--", 0) != 0) return fail(MLBA_CMD_STUPID_USER_ERROR, "unexpected argument " + t);
      const size_t eq = t.find('=');
      if (eq != std::string::npos) {
        flags_[t.substr(2, eq - 2)] = t.substr(eq + 1);
      } else if (i + 1 < tokens.size() && tokens[i + 1].rfind("--", 0) != 0) {
        flags_[t.substr(2)] = tokens[++i];
      } else if (t.rfind("--no", 0) == 0 && known_bool(t.substr(4))) {
        flags_[t.substr(4)] = "false";
      } else {
        flags_[t.substr(2)] = "true";
      }
    }
    for (const auto& kv : flags_)
      if (!known(kv.first)) return fail(MLBA_CMD_STUPID_USER_ERROR, "unknown flag --" + kv.first);
    for (const Command& c : commands_)
      for (const std::string& n : c.names)
        if (n == tokens[0]) return c.callback();
    return fail(MLBA_CMD_STUPID_USER_ERROR, "unknown command " + tokens[0]);
  }

 private:
  void addCommand(std::vector<std::string> names, std::function<int()> cb, std::string help) {
    commands_.push_back(Command{std::move(names), std::move(cb), std::move(help)});
  }

  int fail(int code, const std::string& msg) {
    message_ = msg;
    return code;
  }

  static const std::map<std::string, char>& flag_types() {
    static const std::map<std::string, char> types = {
        {"ba_num_iterations", 'i'}, {"ba_max_time_seconds", 'd'}, {"ba_use_jacobi_scaling", 'b'},
        {"ba_initial_trust_region_radius", 'd'}, {"ba_max_trust_region_radius", 'd'},
        {"ba_min_trust_region_radius", 'd'}, {"ba_function_tolerance", 'd'}, {"ba_gradient_tolerance", 'd'},
        {"ba_parameter_tolerance", 'd'}, {"ba_outlier_rejection_reject_every_n_iters", 'i'},
        {"ba_outlier_rejection_use_reprojection_error", 'b'},
        {"ba_outlier_rejection_max_reprojection_error_px", 'd'}, {"ba_include_visual", 'b'},
        {"ba_include_inertial", 'b'}, {"ba_use_visual_outlier_rejection_solver", 'b'},
        {"ba_fix_landmark_positions", 'b'}, {"ba_fix_vertices", 'b'}, {"ba_fix_ncamera_intrinsics", 'b'},
        {"ba_fix_ncamera_extrinsics_rotation", 'b'}, {"ba_fix_ncamera_extrinsics_translation", 'b'},
        {"ba_min_landmarks_per_frame", 'i'},
        {"elq_min_observers", 'i'}, {"elq_min_observation_angle_deg", 'd'},
        {"elq_max_distance_from_closest_observer", 'd'}, {"elq_min_distance_from_closest_observer", 'd'},
        {"elq_max_reprojection_error_px", 'd'}, {"map_mission", 's'}, {"map_mission_list", 's'}};
    return types;
  }
  static bool known(const std::string& f) { return flag_types().count(f) != 0; }
  static bool known_bool(const std::string& f) {
    auto it = flag_types().find(f);
    return it != flag_types().end() && it->second == 'b';
  }

  // Parses the flag into *out when given; false (with a message) when malformed.
  bool get(const std::string& f, double* out) {
    auto it = flags_.find(f);
    if (it == flags_.end()) return true;
    char* end = nullptr;
    const double v = std::strtod(it->second.c_str(), &end);
    if (end == it->second.c_str() || *end) return (message_ = "bad value for --" + f), false;
    *out = v;
    return true;
  }
  bool get(const std::string& f, int32_t* out) {
    double v = *out;
    if (!get(f, &v)) return false;
    *out = static_cast<int32_t>(v);
    return true;
  }
  bool get_bool(const std::string& f, int32_t* out) {
    auto it = flags_.find(f);
    if (it == flags_.end()) return true;
    const std::string& v = it->second;
    if (v == "true" || v == "1") *out = 1;
    else if (v == "false" || v == "0") *out = 0;
    else return (message_ = "bad value for --" + f), false;
    return true;
  }

  // ViProblemOptions::initFromGFlags and OptimizerPlugin::optimize
  // (vi-optimization-builder.cc:98-150, optimizer-plugin.cc:190-235).
  int optimize(bool inertial) {
    mlba_vi_problem_options o;
    mlba_vi_problem_options_init(&o);
    int32_t include_inertial = inertial ? 1 : 0;
    bool ok = get("ba_num_iterations", &o.solver.max_num_iterations) &&
              get("ba_max_time_seconds", &o.solver.max_solver_time_in_seconds) &&
              get_bool("ba_use_jacobi_scaling", &o.solver.use_jacobi_scaling) &&
              get("ba_initial_trust_region_radius", &o.solver.initial_trust_region_radius) &&
              get("ba_max_trust_region_radius", &o.solver.max_trust_region_radius) &&
              get("ba_min_trust_region_radius", &o.solver.min_trust_region_radius) &&
              get("ba_function_tolerance", &o.solver.function_tolerance) &&
              get("ba_gradient_tolerance", &o.solver.gradient_tolerance) &&
              get("ba_parameter_tolerance", &o.solver.parameter_tolerance) &&
              get("ba_outlier_rejection_reject_every_n_iters", &o.solver.reject_outliers_every_n_iters) &&
              get_bool("ba_outlier_rejection_use_reprojection_error", &o.solver.use_reprojection_error) &&
              get("ba_outlier_rejection_max_reprojection_error_px", &o.solver.max_reprojection_error_px) &&
              get_bool("ba_include_visual", &o.add_visual_constraints) &&
              get_bool("ba_use_visual_outlier_rejection_solver", &o.enable_visual_outlier_rejection) &&
              get_bool("ba_fix_landmark_positions", &o.fix_landmark_positions) &&
              get_bool("ba_fix_vertices", &o.fix_vertices) &&
              get_bool("ba_fix_ncamera_intrinsics", &o.calibration.fix_intrinsics) &&
              get_bool("ba_fix_ncamera_extrinsics_rotation", &o.calibration.fix_extrinsics_rotation) &&
              get_bool("ba_fix_ncamera_extrinsics_translation", &o.calibration.fix_extrinsics_translation) &&
              get("ba_min_landmarks_per_frame", &o.min_landmarks_per_frame) &&
              get("elq_min_observers", &o.landmark_quality.min_observers) &&
              get("elq_min_observation_angle_deg", &o.landmark_quality.min_observation_angle_deg) &&
              get("elq_max_distance_from_closest_observer",
                  &o.landmark_quality.max_distance_from_closest_observer) &&
              get("elq_min_distance_from_closest_observer",
                  &o.landmark_quality.min_distance_from_closest_observer) &&
              get("elq_max_reprojection_error_px", &o.landmark_quality.max_reprojection_error_px) &&
              (!inertial || get_bool("ba_include_inertial", &include_inertial));
    if (!ok) return MLBA_CMD_STUPID_USER_ERROR;
    // --ba_fix_ncamera_intrinsics covers the intrinsics and the distortion
    // (optimization-terms-addition.cc:223-228).
    o.calibration.fix_distortion = o.calibration.fix_intrinsics;
    if (include_inertial)
      return fail(MLBA_CMD_UNKNOWN_ERROR,
                  "inertial terms are evaluated (mlba_inertial_evaluate) but not part of the device solve; "
                  "run optvi with --ba_include_inertial=false or use optv");
    if (!console_ || !console_->selected_map) return fail(MLBA_CMD_STUPID_USER_ERROR, "No map selected.");
    mlba_console_map* m = console_->selected_map;
    const int M = m->desc.num_missions;
    std::vector<uint8_t> missions;
    if (flags_.count("map_mission") && flags_.count("map_mission_list"))
      return fail(MLBA_CMD_STUPID_USER_ERROR, "Please provide only one of --map_mission and --map_mission_list.");
    for (const char* f : {"map_mission", "map_mission_list"}) {
      auto it = flags_.find(f);
      if (it == flags_.end()) continue;
      missions.assign(M, 0);
      std::stringstream ss(it->second);
      for (std::string item; std::getline(ss, item, ',');) {
        char* end = nullptr;
        const long k = std::strtol(item.c_str(), &end, 10);
        if (end == item.c_str() || *end || k < 0 || k >= M)
          return fail(MLBA_CMD_STUPID_USER_ERROR, "The given mission \"" + item + "\" is not valid.");
        missions[k] = 1;
      }
    }
    const int rc = mlba_vi_map_optimize_cameras(&m->desc, &o, missions.empty() ? nullptr : missions.data(),
                                                m->landmark_quality, m->vertices, m->landmarks,
                                                m->camera_q_CI__C_p_CI, m->camera_intrinsics, m->camera_distortion,
                                                m->camera_xi, console_->last_summary);
    if (rc != MLBA_OK) return fail(MLBA_CMD_UNKNOWN_ERROR, mlba_optimizer_last_error());
    return MLBA_CMD_SUCCESS;
  }

  mlba_console* console_;
  std::vector<Command> commands_;
  std::map<std::string, std::string> flags_;
  std::string message_;
  std::vector<std::string> name_list_;  // built once: returned pointers stay valid while the plugin lives
  friend const char* ::mlba_plugin_command_name(const void*, int32_t);
  friend int32_t(::mlba_plugin_num_commands)(const void*);
};

}  // namespace

extern "C" {

void* mlba_plugin_create(mlba_console* console, void* /*plotter*/) { return new OptimizerPlugin(console); }

void mlba_plugin_destroy(void* plugin) { delete static_cast<OptimizerPlugin*>(plugin); }

const char* mlba_plugin_id(const void* plugin) {
  return plugin ? static_cast<const OptimizerPlugin*>(plugin)->id() : "";
}

int32_t mlba_plugin_num_commands(const void* plugin) {
  return plugin ? (int32_t)static_cast<const OptimizerPlugin*>(plugin)->name_list_.size() : 0;
}

const char* mlba_plugin_command_name(const void* plugin, int32_t index) {
  if (!plugin) return "";
  const OptimizerPlugin* p = static_cast<const OptimizerPlugin*>(plugin);
  return (index >= 0 && index < (int32_t)p->name_list_.size()) ? p->name_list_[index].c_str() : "";
}

int32_t mlba_plugin_process_command(void* plugin, const char* line) {
  if (!plugin || !line) return MLBA_CMD_STUPID_USER_ERROR;
  return static_cast<OptimizerPlugin*>(plugin)->process(line);
}

const char* mlba_plugin_last_message(const void* plugin) {
  return plugin ? static_cast<const OptimizerPlugin*>(plugin)->message().c_str() : "";
}

}  // extern "C"
