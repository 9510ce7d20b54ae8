/* C-ABI of the MI355X SDL segment-linking library (libsdl_amd.so).
 *
 * The reference (bsathian/SDL) exposes a C++ class, SDL::Event (Event.h:24-411),
 * plus three process-global geometry singletons loaded from text files
 * (TiltedGeometry.h:33, EndcapGeometry.h:45, ModuleConnectionMap.h:32).  This
 * header is the drop-in boundary that replaces them: plain pointers and sizes,
 * int status codes (0 = ok, < 0 = error, message in sdl_last_error()).  The C++
 * shim in sdl_amd/include/SDL/ re-exposes the reference's class API on top of it.
 *
 * One handle (sdl_event) holds a batch of one or more events and owns one HIP
 * stream; handles are independent.  All device work of a handle is ordered on
 * its stream; the geometry is immutable after loading and may be shared by
 * any number of handles.
 */
#ifndef SDL_AMD_H
#define SDL_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sdl_geom sdl_geom;
typedef struct sdl_event sdl_event;
typedef struct sdl_mlp sdl_mlp;

enum {
    SDL_OK = 0,
    SDL_ERR_ARG = -1,
    SDL_ERR_IO = -2,
    SDL_ERR_HIP = -3,
    SDL_ERR_UNKNOWN_DETID = -4,
    SDL_ERR_DUPLICATE_CONNECTION = -5,
    SDL_ERR_CAPACITY = -6,
    SDL_ERR_STATE = -7
};

/* Stage identifiers of sdl_event_run / export (Event.h:240-300). */
enum {
    SDL_STAGE_MD = 0,   /* createMiniDoublets                   Event.cc:245-305 */
    SDL_STAGE_SG = 1,   /* createSegmentsWithModuleMap          Event.cc:445-522 */
    SDL_STAGE_TP = 2,   /* createTriplets                       Event.cc:524-597 */
    SDL_STAGE_TL = 3,   /* createTrackletsWithModuleMap         Event.cc:611-693 */
    SDL_STAGE_TCP = 4,  /* createTrackCandidatesFromTriplets    Event.cc:874-982 */
    SDL_STAGE_TCL = 5,  /* createTrackCandidatesFromTracklets   Event.cc:984-1062 */
    /* layer-pair family (objects stored and exported as SG, TL and TCL) */
    SDL_STAGE_SG_LAYERS = 6,  /* createSegments         Event.cc:432-443, 793-821 */
    SDL_STAGE_TL_LAYERS = 7,  /* createTracklets        Event.cc:599-609, 755-790 */
    SDL_STAGE_TC_LAYERS = 8,  /* createTrackCandidates  Event.cc:823-872 */
    /* navigation: tracklets (inner SG, SG two hops out), stored as TL */
    SDL_STAGE_TL_NAV = 9,     /* createTrackletsViaNavigation  Event.cc:696-751 */
    /* pseudo mini-doublets, stored as MD: each anchor-side hit with the first
     * hit of the partner sensor */
    SDL_STAGE_MD_PSEUDO = 10  /* createPseudoMiniDoubletsFromAnchorModule  Event.cc:307-430 */
};

/* Last error message of the calling thread. */
const char* sdl_last_error(void);

/* Loads the three geometry files in the reference's formats:
 *   tilted  "detid drdz slope"                 (TiltedGeometry.cc:18-41)
 *   endcap  "detid avgr2 yl sl yh sh c7 c8 cz" (EndcapGeometry.cc:18-61; column 7 is read as the centroid phi)
 *   map     "detid n d1 ... dn"                (ModuleConnectionMap.cc:18-47)
 * and uploads an immutable device copy (module table + CSR connection map) to
 * the current HIP device.  A connection list holding the same detId twice is
 * rejected with SDL_ERR_DUPLICATE_CONNECTION. */
int sdl_geom_load(const char* tiltedPath, const char* endcapPath, const char* mapPath, sdl_geom** out);
void sdl_geom_free(sdl_geom* g);
/* Number of sensor modules known to the geometry. */
int sdl_geom_num_modules(const sdl_geom* g);
/* Binary geometry cache (no counterpart in the reference; it replaces the text
 * parsing of TiltedGeometry::load, EndcapGeometry::load and
 * ModuleConnectionMap::load, TiltedGeometry.cc:18-41, EndcapGeometry.cc:18-61,
 * ModuleConnectionMap.cc:18-47, at start-up).  The cache holds the parsed module
 * table and CSR connection map exactly as the device uses them, with a version,
 * the record size and a checksum; a file that fails any check is rejected with
 * SDL_ERR_IO.  sdl_geom_convert parses the three text files and writes the
 * cache without touching the GPU; sdl_geom_save writes a loaded geometry's
 * cache; sdl_geom_load_binary loads a cache onto the current HIP device;
 * sdl_geom_check_binary validates a cache on the host only. */
int sdl_geom_convert(const char* tiltedPath, const char* endcapPath, const char* mapPath, const char* binaryPath);
int sdl_geom_save(const sdl_geom* g, const char* binaryPath);
int sdl_geom_load_binary(const char* binaryPath, sdl_geom** out);
int sdl_geom_check_binary(const char* binaryPath, int* numModules, long long* numConnections);

/* The geometry from tables held in memory instead of the three files: the
 * content of the reference's singletons after TiltedGeometry::load,
 * EndcapGeometry::load and ModuleConnectionMap::load / add (TiltedGeometry.cc:18-41,
 * EndcapGeometry.cc:18-61, ModuleConnectionMap.cc:18-85).  nT tilted rows
 * (detId, drdz, slope); nE endcap rows (detId and its eight float columns
 * avgr2 yl sl yh sh c7 c8 cz, 8 per row in eCols, column 7 read as the centroid
 * phi); nM connection rows (detId, mCount[i] connected detIds, concatenated in
 * mList).  Same checks and result as sdl_geom_load on files with these rows. */
int sdl_geom_load_tables(int nT, const unsigned* tDet, const float* tDrdz, const float* tSlope, int nE,
                         const unsigned* eDet, const float* eCols, int nM, const unsigned* mDet, const int* mCount,
                         const unsigned* mList, sdl_geom** out);

/* The reference's detId decoding (Module.cc:159-358): out = subdet, side,
 * layer, rod, ring, module, isLower, isInverted, partnerDetId, moduleType
 * (0 PS, 1 2S), moduleLayerType (0 Pixel, 1 Strip).  Host only. */
int sdl_module_decode(unsigned int detId, unsigned int out[11]);

/* Connection map from simulated tracks (no counterpart in the reference, whose
 * map files were produced outside it; the result follows its format,
 * ModuleConnectionMap.cc:36-44, and the sort + unique of
 * ModuleConnectionMap::add, ModuleConnectionMap.cc:49-85).  Track t is the run
 * [trackOffsets[t], trackOffsets[t+1]) of (station, detId) crossings with
 * non-decreasing stations (one station: one layer or disk crossing); every
 * lower module of a station connects to every lower module of the next
 * station of the track (an upper sensor counts as its lower partner).  Writes
 * "detid n d1 ... dn" lines, sorted by detId, each list sorted without
 * duplicates, to mapPath.  With merge != 0 mapPath is first read as
 * ModuleConnectionMap::load reads it (a repeated detId keeps its last line) and
 * the new lists are merged as ModuleConnectionMap::add merges a second file: a
 * detId with new connections gets the sorted union without duplicates, every
 * other line is written back as read.  Host only. */
int sdl_map_build(long long nTracks, const long long* trackOffsets, const int* station, const unsigned* detId,
                  const char* mapPath, int merge);

/* Creates an event handle (batch of events) on the geometry's device. */
int sdl_event_create(sdl_geom* g, sdl_event** out);
void sdl_event_destroy(sdl_event* ev);

/* Clears all hits and objects (keeps device allocations for reuse). */
int sdl_event_reset(sdl_event* ev);

/* Appends hits of the current (last) event from HOST arrays: the buffered form of
 * Event::addHitToModule(Hit(x, y, z, idx), detId) (Event.cc:120-139).  idx must be
 * unique within an event (objects are matched by hit index, Hit.cc:175-182). */
int sdl_event_add_hits(sdl_event* ev, int n, const float* x, const float* y, const float* z, const int* idx,
                       const unsigned int* detId);
/* Starts a new event in the batch (the next add_hits call fills it). */
int sdl_event_next(sdl_event* ev);

/* Sets the whole batch from DEVICE-resident arrays: hits of event e occupy
 * [offsets[e], offsets[e+1]) (offsets is a HOST array of nevents+1 entries).
 * Every detId must be known to the geometry (else SDL_ERR_UNKNOWN_DETID). */
int sdl_event_set_hits_device(sdl_event* ev, int nevents, const long long* offsets, const float* dx, const float* dy,
                              const float* dz, const int* didx, const unsigned int* ddetId);

/* Runs one stage on every event of the batch.  Each stage reads the objects of
 * the stage before it, so stages run in order (else SDL_ERR_STATE): the
 * module-map chain MD (or MD_PSEUDO), SG, TP, TL, TCP, TCL; SG_LAYERS after MD
 * or MD_PSEUDO; TL_LAYERS after
 * SG or SG_LAYERS; TL_NAV after SG, SG_LAYERS or TP; TC_LAYERS after TL,
 * TL_LAYERS or TL_NAV; TCL also after TL_NAV over module-map segments (the
 * module-map joins TP, TCP and TCL assume module-map inputs and do not follow a
 * layer-pair stage).
 * The layer-pair stages store their objects as the SG, TL and TCL objects of
 * the batch.  algo takes the reference's Algo.h values (Algo.h:6-40): 0 =
 * Default for every stage, 1 = AllComb (MiniDoublet.cc:223-226,
 * Segment.cc:313-335, Triplet.cc:45-48, TrackCandidate.cc:138-141: every
 * candidate accepted, the objects keeping their constructors' values); for the
 * three tracklet stages 1 = DefaultNm1_TLAlgo (Tracklet.cc:127-131: the
 * default cuts with the last one, dBeta, waived, Tracklet.cc:1513-1536,
 * 1990-2014, 2636-2653) and 2 = AllComb (Tracklet.cc:139-142).  TCs from
 * triplets ignore algo, as the reference does (Event.cc:874-982: the bridges
 * always run the default segment and tracklet algorithms), and pseudo
 * mini-doublets always run AllComb (Event.cc:353).  Any other value is
 * rejected with SDL_ERR_ARG (the reference warns and stores nothing).  TCs from
 * tracklets may follow the module-map tracklets directly.  TCs from triplets
 * after AllComb segments or triplets are refused (SDL_ERR_STATE): their
 * bridges are default segments and tracklets built on objects the device does
 * not store.  sdl_event_run_chain always runs the defaults. */
int sdl_event_run(sdl_event* ev, int stage, int algo);
/* Runs MD, SG, TP, TL, TCP and TCL back to back.  After a first call has
 * measured the batch's object counts, later calls size their buffers from
 * them (scaled by the hit count, 25 % headroom) and enqueue the whole chain
 * with one host round trip; if the kernels find a buffer too small, the
 * chain resumes from that stage with buffers sized exactly (same results). */
int sdl_event_run_chain(sdl_event* ev);
/* Waits for the handle's stream. */
int sdl_event_sync(sdl_event* ev);

/* Counters of event `e` (-1 = whole batch): out[0..4] candidates considered for
 * MD, SG, TP, TL, TC; out[5..9] objects formed (Event.cc:1064-1167). */
int sdl_event_counts(sdl_event* ev, int e, unsigned long long out[10]);
/* Per-layer counters of event e (Event.cc:1169-1344): out[11*k + j] for
 * k = 0..9 (as above) and j = 0..5 barrel layers 1-6, 6..10 endcap layers 1-5. */
int sdl_event_layer_counts(sdl_event* ev, int e, unsigned long long out[110]);

/* Exports the objects of one stage of event e as hit-index tuples
 * (MD 2, SG 4, TP 8, TL 8, TCP 16, TCL 16 ints per object) into out (cap ints).
 * *n receives the number of objects; returns SDL_ERR_CAPACITY if cap is short. */
int sdl_event_export(sdl_event* ev, int e, int stage, int* out, long long cap, long long* n);
/* sdl_event_export without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  The tuples are gathered in
 * a static 64 MiB device array per GPU (16777216 ints) instead of the event's
 * growable export buffer; an export larger than that returns SDL_ERR_CAPACITY.
 * The first call on a GPU looks the array up (hipGetSymbolAddress).  Calls
 * on one GPU share the array: calls that overlap in time on one GPU race. */
int sdl_event_export_noalloc(sdl_event* ev, int e, int stage, int* out, long long cap, long long* n);

/* Per-object values of event e's stored MDs, SGs, TLs or TCs from tracklets
 * (stage SDL_STAGE_MD, _SG, _TL or _TCL; also the objects of the layer-pair, navigation and pseudo-MD
 * stages): what the reference keeps on each stored object, its typed getters
 * and setRecoVars maps (MiniDoublet.h:85-97 / MiniDoublet.cc:113-197,
 * Segment.h:128-149 / Segment.cc:104-284, Tracklet.cc:43-98,
 * TrackletBase.cc:50-60), which the reference's users read through the
 * objects' getters.  sdl_object_var_count(stage) floats per object (MD 7,
 * SG 22, TL 114: six typed values, then the 108 keys of the tracklet's
 * reco-variable map, Tracklet.cc:144-296,902-2665,3181-3324, in std::map order,
 * named reco_<key>; TCL 2: dR, innerR of a TC made from tracklets,
 * TrackCandidate.cc:206-207; 0 for the other stages, TCs from triplets run no
 * algorithm and record nothing), named by sdl_object_var_names(stage)
 * (comma-separated), in the object order of sdl_event_export.  Fields a path
 * never sets keep the reference's constructor value 0; reco variables a path
 * never sets are -999, or NaN where the map has no entry (pseudo MDs' miniCut,
 * the ten keys only the barrel tracklet algorithm writes).
 * Returns SDL_ERR_CAPACITY if cap (floats) is short; out = NULL returns *n. */
int sdl_object_var_count(int stage);
const char* sdl_object_var_names(int stage);
int sdl_event_export_vars(sdl_event* ev, int e, int stage, float* out, long long cap, long long* n);

/* Number of events in the batch. */
int sdl_event_num_events(const sdl_event* ev);

/* The HIP stream of the handle (hipStream_t), for callers that order their own work after it. */
void* sdl_event_stream(sdl_event* ev);

/* Diagnostics with no counterpart in the reference: when timing is on, the
 * hit-preparation, MD, SG and TL kernels of each run are bracketed with HIP
 * events on the handle's stream; sdl_event_kernel_ms writes their durations
 * (ms, in that order, n <= 4) for the last run. */
int sdl_event_set_timing(sdl_event* ev, int on);
int sdl_event_kernel_ms(sdl_event* ev, double* out, int n);
/* With no counterpart in the reference: whether run_chain may capture its
 * speculative form into a HIP graph and replay it (default on). */
int sdl_event_set_graphs(sdl_event* ev, int on);
/* With no counterpart in the reference: counts since the handle was created,
 * out[0] stage runs (sdl_event_run or sdl_event_run_chain calls), out[1] runs
 * replayed from a captured graph, out[2] graph captures, out[3] exact reruns
 * (a speculative run whose predicted sizes fell short; the rerun sizes each
 * stage from its read-back totals and a captured graph is captured anew). */
int sdl_event_run_stats(sdl_event* ev, long long out[4]);
/* Tuning knob with no counterpart in the reference: the headroom factor of the
 * predicted buffer sizes for the stages >= fromStage (default 1.25 for all).
 * A factor below 1 forces the resume path; the tests use it to check it. */
int sdl_event_set_headroom(sdl_event* ev, int fromStage, double factor);
/* Autotuner with no counterpart in the reference.  Launch shapes of the main
 * kernels (values in this order): md_block, sg_block, tl1_block, tl2_block,
 * tl_expand_block (block sizes 64, 128 or 256 of the mini-doublet, segment,
 * first- and second-stage tracklet and candidate-expansion kernels) and
 * scored_waves (waves per block of the learned proposal's first stage: 2, 4
 * or 8).
 * sdl_event_autotune times run_chain on the handle's current batch (median of
 * `reps` runs per setting), tries each knob's values in turn and keeps the
 * fastest (a change must gain 1 %); the timings and choice are written to
 * `report` as JSON.  Object sets do not depend on the launch shapes.
 * get/set return or restore a choice (n = 6 values; get returns 6). */
int sdl_event_autotune(sdl_event* ev, int reps, char* report, int cap);
int sdl_event_get_tuning(sdl_event* ev, int* values, int n);
int sdl_event_set_tuning(sdl_event* ev, const int* values, int n);

/* ---------------------------------------------------------------- learned proposal
 * No counterpart in the reference.  A 24-64-64-64-1 MLP scores the tracklet
 * stage's (inner segment, outer-inner mini-doublet) pairs that passed the
 * first-stage cuts, before the second-stage candidates are formed; pairs scored
 * below the threshold are dropped, every object that is kept still passes all
 * exact cuts.  Inference runs in BF16 MFMA with
 * FP32 accumulation; training runs FP32 forward and backward kernels with Adam.
 * Parameters: float32, layout W1[64][24], b1[64], W2[64][64], b2[64],
 * W3[64][64], b3[64], w4[64], b4 (sdl_mlp_num_params() values). */
int sdl_mlp_create(sdl_mlp** out);
void sdl_mlp_free(sdl_mlp* m);
int sdl_mlp_num_params(void);
int sdl_mlp_set_params(sdl_mlp* m, const float* params, int n);
int sdl_mlp_get_params(sdl_mlp* m, float* params, int n);
/* Gradient of the mean weighted binary cross entropy (positives weighted by
 * posWeight) over n DEVICE rows dX (n x 24) with labels dY, without an update. */
int sdl_mlp_gradient(sdl_mlp* m, const float* dX, const float* dY, long long n, float posWeight, float* grad,
                     float* loss);
/* `steps` Adam steps on consecutive batches of rows dIdx[start + k*batch ...]
 * (device indices; NULL: the rows themselves); *lastLoss: the last batch's loss. */
int sdl_mlp_train(sdl_mlp* m, const float* dX, const float* dY, const int* dIdx, long long start, int batch, int steps,
                  float lr, float posWeight, float* lastLoss);
/* Logits of n DEVICE feature rows through the BF16 MFMA inference path. */
int sdl_mlp_forward(sdl_mlp* m, const float* dX, long long n, float* dLogit);
/* Uses m to drop tracklet pairs whose score is below threshold (0 < t < 1);
 * m = NULL restores the plain chain. */
int sdl_event_set_proposal(sdl_event* ev, sdl_mlp* m, float threshold);
/* Training rows from the last plain chain: features (24 per pair, DEVICE dX)
 * and labels (DEVICE dY: 1 if the pair yields a tracklet, 0 if it passed the
 * first-stage cuts but yields none, -1 if it failed them and is never scored)
 * of every pair; with dX = NULL only *n (the number of pairs) is returned. */
int sdl_event_tl_pair_rows(sdl_event* ev, float* dX, float* dY, long long cap, long long* n);
/* Diagnostics of the last tracklet stage: out[0] (inner segment, MD) pairs,
 * out[1] candidates that reached the second-stage cuts, out[2] tracklets,
 * out[3] pairs scored by the learned proposal (0 without one). */
int sdl_event_tl_work(sdl_event* ev, long long out[4]);

#ifdef __cplusplus
}
#endif

#endif
