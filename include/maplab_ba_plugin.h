/*
 * maplab_ba_plugin.h -- console-plugin boundary of the map-optimization
 * plugin (console-plugins/map-optimization-plugin/src/optimizer-plugin.cc:
 * 27-55, 190-235): plugin id "map-optimization", commands optimize_visual /
 * optv and optimize_visual_inertial / optvi run synchronously, options from
 * the reference's flag names (--ba_*, --elq_*, --map_mission,
 * --map_mission_list) given on the command line. The console is the small
 * host structure below: it holds the selected map (the VIMapManager map
 * under write access in the reference).
 *
 * The entry points are C functions with C types. maplab's console resolves
 * createConsolePlugin / destroyConsolePlugin with C++ types
 * (console-plugin-base-with-plotter.h:43-56, maplab-console.cc:129-138); this
 * library does not export those names, so it cannot be mistaken for a
 * plugin of that ABI. A C++ shim compiled against maplab exports them and
 * forwards to mlba_plugin_create / mlba_plugin_destroy (shown in the
 * integration notes).
 */
#ifndef MAPLAB_BA_PLUGIN_H_
#define MAPLAB_BA_PLUGIN_H_

#include <stdint.h>

#include "maplab_ba.h"

#ifdef __cplusplus
extern "C" {
#endif

/* common::CommandStatus (console-common/command-registerer.h:19-24). */
#define MLBA_CMD_SUCCESS 0
#define MLBA_CMD_UNKNOWN_ERROR 1
#define MLBA_CMD_STUPID_USER_ERROR 2

/* The selected map: its description, and the state buffers the command
 * writes back into (7 P vertices, 3 L landmarks, L landmark qualities,
 * MLBA_QUALITY_* of maplab_ba_optimizer.h; qualities nullable). */
typedef struct mlba_console_map {
  mlba_problem_desc desc;
  double* vertices;
  double* landmarks;
  uint8_t* landmark_quality;
  /* Camera blocks written back by a calibration refinement (7C, 4C, 4C, C;
   * nullable: not written back). */
  double* camera_q_CI__C_p_CI;
  double* camera_intrinsics;
  double* camera_distortion;
  double* camera_xi;
} mlba_console_map;

/* The console a plugin is created for. selected_map NULL = no map selected. */
typedef struct mlba_console {
  mlba_console_map* selected_map;
  mlba_summary* last_summary;     /* nullable: receives the summary of the last optimization */
} mlba_console;

/* Plugin lifetime: createConsolePlugin / destroyConsolePlugin
 * (console-plugin-base-with-plotter.h:43-56) with C types; plotter unused,
 * may be NULL. */
void* mlba_plugin_create(mlba_console* console, void* plotter);
void mlba_plugin_destroy(void* plugin);

/* The plugin id ("map-optimization") and its command names; the returned
 * strings stay valid until mlba_plugin_destroy. */
const char* mlba_plugin_id(const void* plugin);
int32_t mlba_plugin_num_commands(const void* plugin);
const char* mlba_plugin_command_name(const void* plugin, int32_t index);

/* Runs one console line, "<command> [--flag=value | --flag value | --flag |
 * --noflag]...", and returns its CommandStatus; unknown commands and flags
 * are user errors. The reason of a failure is in mlba_plugin_last_message. */
int32_t mlba_plugin_process_command(void* plugin, const char* line);
const char* mlba_plugin_last_message(const void* plugin);

#ifdef __cplusplus
}
#endif
#endif /* MAPLAB_BA_PLUGIN_H_ */
