/*
 * rexlib_hip_plugin.h — the entry points of librexlib-hip.so, the MI355X
 * backend as a rexlib plugin (rexlib_amd/host/hip_plugin.cpp).
 *
 * These are the two symbols rexlib's loader looks up in a plugin's shared
 * object (include/rexlib/core/plugin_definition.hpp:26-50, expanded from
 * REXLIB_DEFINE_PLUGIN; loader src/core/plugin_manager.cpp, checks
 * src/core/plugin_compatibility_check.cpp). Everything else the plugin
 * provides is reached through the C++ objects the second one returns:
 * rexlib::plugin::register_at adds the "hip" device backend and the program
 * builders of insert_slices, insert_weighted_slices, insert_ctf_slices,
 * rfft, irfft, fill and copy, which run the kernels of librexlib_hip.so
 * (include/rexlib_hip.h).
 */
#ifndef REXLIB_HIP_PLUGIN_H
#define REXLIB_HIP_PLUGIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* rexlib::plugin_compatibility (include/rexlib/core/plugin_compatibility.hpp:23-62):
 * plain data of a fixed layout, readable whatever the C++ ABI. */
typedef struct rexlib_plugin_compatibility_record {
	uint32_t size;                 /* bytes of the record the plugin was built with (32) */
	uint32_t rexlib_version_major; /* rexlib release the plugin was built against: 0.1.0 */
	uint32_t rexlib_version_minor;
	uint32_t rexlib_version_patch;
	uint32_t abi_family;           /* 1: Itanium C++ ABI */
	uint32_t standard_library;     /* 1: libstdc++ */
	uint32_t standard_library_abi; /* _GLIBCXX_USE_CXX11_ABI */
	uint32_t runtime;              /* 0: no MSVC runtime */
} rexlib_plugin_compatibility_record;

/* Replaces the plugins' rexlib_get_plugin_compatibility
 * (include/rexlib/core/plugin_definition.hpp:28-34): the record of this
 * build, read by the loader before anything else. */
const rexlib_plugin_compatibility_record *rexlib_get_plugin_compatibility(void);

/* Replaces the plugins' rexlib_get_plugin (include/rexlib/core/plugin_definition.hpp:35-46):
 * a `const rexlib::plugin *` to a function-local static constructed on the
 * first call, or null when its constructor throws. Constructing it makes no
 * HIP call, so it loads on a host without a GPU. */
const void *rexlib_get_plugin(void);

#ifdef __cplusplus
}
#endif

#endif
