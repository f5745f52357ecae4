/*
 * maplab_ba_vimap.h -- maplab map folders to the visual BA problem and back,
 * for a C / C++ host (the counterpart of maplab_amd/vimap.py).
 *
 * Reads what vi_map::serialization::loadMapFromFolder reads for the BA
 * (map-structure/vi-map/include/vi-map/vi-map-serialization.h:54-78,
 * vi-map-serialization.cc:228-251): `vi_map/vi_map_metadata` and the
 * (optionally gzip-compressed) VIMap protobuf files it lists (missions,
 * vertices, edges, landmark index), and the NCAMERA sensors of
 * `vi_map/sensors.yaml`; and builds the problem the way
 * OptimizationStateBuffer::importStatesOfMissions builds its blocks
 * (algorithms/map-optimization/src/optimization-state-buffer.cc:153-263):
 * vertex blocks [q_IM xyzw (Hamilton q_M_I, w >= 0) | p_MI], baseframe
 * blocks JPL q_GM, camera blocks JPL q_CI from the NCamera's T_B_C; vertices
 * along each mission's VIWLS chain from its root (the root is the constant
 * gauge vertex), landmarks from the vertices' landmark stores, one
 * observation per keypoint whose landmark id is known. Saving writes the
 * optimized vertex poses and landmark positions (and, optionally, landmark
 * qualities) back into the same protobuf messages; every other field,
 * unknown ones included, is copied unchanged.
 *
 * Protobuf wire format and gzip are handled by the library (zlib); no
 * protobuf or YAML library is needed.
 */
#ifndef MAPLAB_BA_VIMAP_H_
#define MAPLAB_BA_VIMAP_H_

#include <stdint.h>

#include "maplab_ba.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlba_vimap mlba_vimap;

/* Loads the map folder (the folder that contains vi_map/). */
int32_t mlba_vimap_load(const char* folder, mlba_vimap** out_map);
void mlba_vimap_free(mlba_vimap* map);

/* The problem: every pointer of desc points into arrays owned by the map
 * (valid until mlba_vimap_free). camera_model / camera_xi are NULL when
 * every camera is pinhole. */
int32_t mlba_vimap_describe(const mlba_vimap* map, mlba_problem_desc* desc);

/* Landmark.quality of every landmark (vi_map.proto Landmark.Quality:
 * 0 unknown, 1 bad, 2 good), L entries. */
int32_t mlba_vimap_landmark_quality(const mlba_vimap* map, uint8_t* quality);

/* Writes the vertex poses (7 P, the describe layout), landmark positions
 * (3 L, p_B) and, when landmark_quality is not NULL, the landmark qualities
 * that are 1 or 2 into the map's messages and saves every VIMap file into
 * `folder`/vi_map/ (the metadata and sensors.yaml are copied), each file
 * compressed as it was read. folder may be the map's own folder. */
int32_t mlba_vimap_save(const mlba_vimap* map, const char* folder, const double* vertices,
                        const double* landmarks, const uint8_t* landmark_quality);

/* Reason of the last failed mlba_vimap_* call of this thread. */
const char* mlba_vimap_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MAPLAB_BA_VIMAP_H_ */
