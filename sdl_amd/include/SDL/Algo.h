// Algorithm enums of the SDL API, same names and values as the reference's Algo.h:6-49.
#ifndef SDL_AMD_ALGO_H
#define SDL_AMD_ALGO_H

namespace SDL {
enum MDAlgo { Default_MDAlgo = 0, AllComb_MDAlgo };
enum SGAlgo { Default_SGAlgo = 0, AllComb_SGAlgo };
enum TLAlgo { Default_TLAlgo = 0, DefaultNm1_TLAlgo, AllComb_TLAlgo };
enum TPAlgo { Default_TPAlgo = 0, AllComb_TPAlgo };
enum TCAlgo { Default_TCAlgo = 0, AllComb_TCAlgo };
enum SDLMode { All_SDLAlgo = 0, Tracklet_SDLAlgo, Triplet_SDLAlgo };
// PrintUtil.h:33-43
enum LogLevel { Log_Nothing = 0, Log_Info = 1, Log_Warning = 2, Log_Error = 3, Log_Critical = 4, Log_Debug = 5, Log_Debug2 = 6, Log_Debug3 = 7 };
}  // namespace SDL

#endif
