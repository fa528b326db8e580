/* Public C ABI of the MI355X NCC find path: the subset of mvkit's C ABI a
 * caller of the NCC operators binds, with the reference's names, types,
 * ownership and error conventions.  Each declaration cites the reference
 * header it replaces (paths relative to the mvkit source tree).
 *
 * Conventions (mvkit include/mvkit/c/mv_tuple.h:18-32, mv_api.h:57-87):
 *  - every function returns an MvStatus; nothing throws across the ABI;
 *  - inputs are borrowed for the call, outputs are new owned tuples that the
 *    caller releases with mv_destroy_tuple; all outputs of one call are
 *    committed together or not at all;
 *  - on failure the thread-local last error holds status, code, operation,
 *    parameter, parameter index and message. */
#ifndef MVKIT_AMD_MV_NCC_H
#define MVKIT_AMD_MV_NCC_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define MV_API __attribute__((visibility("default")))
#else
#define MV_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* include/mvkit/c/mv_status.h:10-24 */
typedef int32_t MvStatus;
#define MV_STATUS_OK ((int32_t)0)
#define MV_STATUS_INVALID_ARGUMENT ((int32_t)-1)
#define MV_STATUS_OUT_OF_RANGE ((int32_t)-2)
#define MV_STATUS_TYPE_MISMATCH ((int32_t)-3)
#define MV_STATUS_NOT_FOUND ((int32_t)-4)
#define MV_STATUS_OUT_OF_MEMORY ((int32_t)-5)
#define MV_STATUS_UNSUPPORTED ((int32_t)-6)
#define MV_STATUS_INVALID_STATE ((int32_t)-7)
#define MV_STATUS_IO_ERROR ((int32_t)-100)
#define MV_STATUS_DECODE_ERROR ((int32_t)-101)
#define MV_STATUS_ENCODE_ERROR ((int32_t)-102)
#define MV_STATUS_LIMIT_EXCEEDED ((int32_t)-103)
#define MV_STATUS_INTERNAL_ERROR ((int32_t)-1000)

/* include/mvkit/c/mv_string.h:10-13 */
typedef struct MvStringView { const char* data; size_t size; } MvStringView;

/* include/mvkit/c/mv_value.h:12-20 */
typedef struct MvHandleImpl* MvHandle;
typedef int32_t MvValueKind;
#define MV_VALUE_EMPTY ((int32_t)0)
#define MV_VALUE_INT64 ((int32_t)1)
#define MV_VALUE_REAL64 ((int32_t)2)
#define MV_VALUE_STRING ((int32_t)4)
#define MV_VALUE_HANDLE ((int32_t)16)

/* include/mvkit/c/mv_tuple.h:18; mv_iconic.h:20 */
typedef struct MvTupleImpl* MvTuple;
typedef struct MvObjectImpl* MvObject;

/* include/mvkit/c/mv_api.h:57-75 */
typedef int32_t MvErrorCode;
#define MV_ERROR_UNSPECIFIED ((int32_t)0)
#define MV_ERROR_NULL_ARGUMENT ((int32_t)1)
#define MV_ERROR_INVALID_CARDINALITY ((int32_t)2)
#define MV_ERROR_INVALID_TYPE ((int32_t)3)
#define MV_ERROR_INVALID_VALUE ((int32_t)4)
#define MV_ERROR_INDEX_OUT_OF_RANGE ((int32_t)5)
#define MV_ERROR_MISSING_ITEM ((int32_t)6)
#define MV_ERROR_UNSUPPORTED_VALUE ((int32_t)7)
#define MV_ERROR_INVALID_OBJECT_STATE ((int32_t)8)
#define MV_ERROR_IO_FAILURE ((int32_t)9)
#define MV_ERROR_INTERNAL_FAILURE ((int32_t)10)
#define MV_ERROR_PARAMETER_INDEX_NONE ((int64_t)-1)
typedef struct MvErrorInfo {
  MvStatus status;
  MvErrorCode code;
  MvStringView operation;
  MvStringView parameter;
  int64_t parameter_index;
  MvStringView message;
} MvErrorInfo;

/* include/mvkit/c/mv_api.h:75, 87 */
MV_API MvStatus mv_get_last_error(MvStringView* error);
MV_API MvStatus mv_get_last_error_info(MvErrorInfo* error);

/* Control tuples: include/mvkit/c/mv_tuple.h:42-228 */
MV_API MvStatus mv_create_tuple(MvTuple* tuple, size_t length);
MV_API MvStatus mv_copy_tuple(MvTuple tuple, MvTuple* clone);
MV_API MvStatus mv_destroy_tuple(MvTuple tuple);
MV_API MvStatus mv_length_tuple(MvTuple tuple, size_t* size);
MV_API MvStatus mv_get_type(MvTuple tuple, size_t index, MvValueKind* kind);
MV_API MvStatus mv_get_i(MvTuple tuple, size_t index, int64_t* value);
MV_API MvStatus mv_get_d(MvTuple tuple, size_t index, double* value);
MV_API MvStatus mv_get_s(MvTuple tuple, size_t index, MvStringView* value);
MV_API MvStatus mv_get_h(MvTuple tuple, size_t index, MvHandle* handle);
MV_API MvStatus mv_resize_tuple(MvTuple tuple, size_t length);
MV_API MvStatus mv_set_i(MvTuple tuple, int64_t value, size_t index);
MV_API MvStatus mv_set_d(MvTuple tuple, double value, size_t index);
MV_API MvStatus mv_set_s(MvTuple tuple, MvStringView value, size_t index);
MV_API MvStatus mv_set_h(MvTuple tuple, MvHandle value, size_t index);

/* Iconic objects: include/mvkit/c/mv_iconic.h:34-40; mv_image.h:45, 78.
 * Type is "byte" or "uint2"; PixelPointer is the integer address of a
 * row-major plane with stride == width.  gen_image1 copies the pixels;
 * gen_image1_extern borrows them and calls ClearProc (integer address of a
 * void(*)(void*), or 0) with the pointer when the object is cleared. */
MV_API MvStatus mv_gen_empty_obj(MvObject* objects);
MV_API MvStatus mv_clear_obj(MvObject objects);
MV_API MvStatus mv_gen_image1(MvObject* Image, MvTuple Type, MvTuple Width, MvTuple Height, MvTuple PixelPointer);
MV_API MvStatus mv_gen_image1_extern(MvObject* Image, MvTuple Type, MvTuple Width, MvTuple Height,
                                     MvTuple PixelPointer, MvTuple ClearProc);

/* Regions and image domains, for reduced-domain (ROI) templates and searches.
 * mv_gen_region_runs: include/mvkit/c/mv_region.h:40 (src/c_api/region_api.cpp:31-78);
 *   one region from runs (row, first column, last column; integers, or reals
 *   rounded by +0.5 and truncation), sorted and merged.
 * mv_get_region_runs: mv_region.h:54; the canonical runs of one region.
 * mv_gen_rectangle1: mv_region.h:71 (region_control_api.cpp:102-113); one
 *   rectangle per control position, corners rounded to the nearest integer.
 * mv_reduce_domain: mv_image_processing.h:166 (image_processing_api.cpp:239-252,
 *   core/image.cpp:267-320); the image with its domain intersected with the
 *   first region clipped to the image.
 * mv_full_domain: mv_image_processing.h:176; the image with its full domain.
 * mv_get_domain: mv_image_processing.h:155; the image's domain as a region. */
MV_API MvStatus mv_gen_region_runs(MvObject* Region, MvTuple Row, MvTuple ColumnBegin, MvTuple ColumnEnd);
MV_API MvStatus mv_get_region_runs(MvObject Region, MvTuple* Row, MvTuple* ColumnBegin, MvTuple* ColumnEnd);
MV_API MvStatus mv_gen_rectangle1(MvObject* Rectangle, MvTuple Row1, MvTuple Column1, MvTuple Row2, MvTuple Column2);
MV_API MvStatus mv_reduce_domain(MvObject Image, MvObject Region, MvObject* ImageReduced);
MV_API MvStatus mv_full_domain(MvObject Image, MvObject* ImageFull);
MV_API MvStatus mv_get_domain(MvObject Image, MvObject* Domain);

/* Handles: include/mvkit/c/mv_handle.h:40.  Invalidates every alias. */
MV_API MvStatus mv_clear_handle(MvTuple Handle);
/* include/mvkit/c/mv_handle.h:52 (src/c_api/handle_api.cpp:67-100): the
 * semantic type of a control tuple as a one-string tuple: "integer", "real",
 * "string", "ncc_model" or "serialized_item" (live handles of one kind),
 * "handle" (cleared or mixed handles), "any" (empty or mixed). */
MV_API MvStatus mv_tuple_sem_type(MvTuple T, MvTuple* SemType);
/* Serialized items: include/mvkit/c/mv_handle.h:63-97
 * (src/c_api/serialization_api.cpp:51-141) for NCC model handles.
 * serialize_handle turns a model into an item holding its MVKIT_NCC 1 bytes;
 * deserialize_handle rebuilds a model from an item; the _ptr pair copies raw
 * bytes into an item (Copy must be "true") and exposes an item's bytes, valid
 * as long as the item's handle is. */
MV_API MvStatus mv_serialize_handle(MvTuple Handle, MvTuple* SerializedItem);
MV_API MvStatus mv_deserialize_handle(MvTuple SerializedItem, MvTuple* Handle);
MV_API MvStatus mv_create_serialized_item_ptr(MvTuple Pointer, MvTuple Size, MvTuple Copy,
                                              MvTuple* SerializedItemHandle);
MV_API MvStatus mv_get_serialized_item_ptr(MvTuple SerializedItemHandle, MvTuple* Pointer, MvTuple* Size);

/* NCC operators: include/mvkit/c/mv_ncc_matching.h:30-123 */
MV_API MvStatus mv_create_ncc_model(MvObject Template, MvTuple NumLevels, MvTuple AngleStart, MvTuple AngleExtent,
                                    MvTuple AngleStep, MvTuple Metric, MvTuple* ModelID);
MV_API MvStatus mv_get_ncc_model_params(MvTuple ModelID, MvTuple* NumLevels, MvTuple* AngleStart,
                                        MvTuple* AngleExtent, MvTuple* AngleStep, MvTuple* Metric);
MV_API MvStatus mv_find_ncc_model(MvObject Image, MvTuple ModelID, MvTuple AngleStart, MvTuple AngleExtent,
                                  MvTuple MinScore, MvTuple NumMatches, MvTuple MaxOverlap, MvTuple SubPixel,
                                  MvTuple NumLevels, MvTuple* Row, MvTuple* Column, MvTuple* Angle, MvTuple* Score);
/* include/mvkit/c/mv_ncc_matching.h:84, 94 (src/c_api/ncc_matching_api.cpp:97-109,
 * src/ncc_matching/persist.cpp:14-84): the model as MVKIT_NCC 1 text (the
 * level-0 template with its mask and centroid, checksummed); reading
 * validates the file and rebuilds the model exactly as create does. */
MV_API MvStatus mv_write_ncc_model(MvTuple ModelID, MvTuple FileName);
MV_API MvStatus mv_read_ncc_model(MvTuple FileName, MvTuple* ModelID);
/* include/mvkit/c/mv_ncc_matching.h:119-123 (src/c_api/ncc_matching_api.cpp:191-203):
 * several models in one image.  ModelIDs holds 1..4096 handles; every other
 * control one value or one per model (NumLevels: highest, a shared pair, one
 * highest per model or a pair per model).  One-value NumMatches / MaxOverlap
 * apply to all models together, per-model values to each model alone, with
 * per-model quotas taken before the cross-model overlap.  Model receives the
 * index of each match's model in ModelIDs. */
MV_API MvStatus mv_find_ncc_models(MvObject Image, MvTuple ModelIDs, MvTuple AngleStart, MvTuple AngleExtent,
                                   MvTuple MinScore, MvTuple NumMatches, MvTuple MaxOverlap, MvTuple SubPixel,
                                   MvTuple NumLevels, MvTuple* Row, MvTuple* Column, MvTuple* Angle, MvTuple* Score,
                                   MvTuple* Model);

#ifdef __cplusplus
}
#endif
#endif
