/*
 * siminf_builtin_models.h -- the reference's built-in models as
 * compile-time plugins of the solver kernels: SIR, SEIR, SIS, SISe, SISe3,
 * SISe_sp, SISe3_sp (the files under src/models) and the code mparse
 * emits for the README's SIR model
 * (S -> beta*S*I/N -> I, I -> gamma*I -> R, N <- S+I+R).  Each model's
 * kernels are instantiated in a translation unit of its own
 * (builtin_*.hip), which exports the model's SimInf_hip_model record.
 */
#pragma once

#include "siminf_kernels.h"

namespace siminf {

struct ModelSIR {
    static constexpr int NC = 3, NT = 2, ND = 0, NLD = 2, NGD = 0;
    static constexpr bool PTS_NOOP = true;   /* pts never changes anything */
    static constexpr int S[NT][NC] =
        {{-1, 1, 0}, {0, -1, 1}};
    static constexpr int G[NT][NT] =
        {{1, 1}, {1, 1}};
    template <int J> SIMINF_MODEL_FN double rate(SIMINF_RATE_ARGS)
    {
        if constexpr (J == 0) return SIR_S_to_I(u, v, ld, gd, t);
        else return SIR_I_to_R(u, v, ld, gd, t);
    }
    SIMINF_MODEL_FN int pts(SIMINF_PTS_ARGS)
    {
        return siminf_pts_noop(vn, u, v, ld, gd, node, t);
    }
};

struct ModelSEIR {
    static constexpr int NC = 4, NT = 3, ND = 0, NLD = 3, NGD = 0;
    static constexpr bool PTS_NOOP = true;   /* pts never changes anything */
    static constexpr int S[NT][NC] =
        {{-1, 1, 0, 0}, {0, -1, 1, 0}, {0, 0, -1, 1}};
    static constexpr int G[NT][NT] =
        {{1, 1, 1}, {1, 1, 1}, {1, 1, 1}};
    template <int J> SIMINF_MODEL_FN double rate(SIMINF_RATE_ARGS)
    {
        if constexpr (J == 0) return SEIR_S_to_E(u, v, ld, gd, t);
        else if constexpr (J == 1) return SEIR_E_to_I(u, v, ld, gd, t);
        else return SEIR_I_to_R(u, v, ld, gd, t);
    }
    SIMINF_MODEL_FN int pts(SIMINF_PTS_ARGS)
    {
        return siminf_pts_noop(vn, u, v, ld, gd, node, t);
    }
};

struct ModelSIS {
    static constexpr int NC = 2, NT = 2, ND = 0, NLD = 2, NGD = 0;
    static constexpr bool PTS_NOOP = true;   /* pts never changes anything */
    static constexpr int S[NT][NC] =
        {{-1, 1}, {1, -1}};
    static constexpr int G[NT][NT] =
        {{1, 1}, {1, 1}};
    template <int J> SIMINF_MODEL_FN double rate(SIMINF_RATE_ARGS)
    {
        if constexpr (J == 0) return SIS_S_to_I(u, v, ld, gd, t);
        else return SIS_I_to_S(u, v, ld, gd, t);
    }
    SIMINF_MODEL_FN int pts(SIMINF_PTS_ARGS)
    {
        return siminf_pts_noop(vn, u, v, ld, gd, node, t);
    }
};

struct ModelSISe {
    static constexpr int NC = 2, NT = 2, ND = 1, NLD = 4, NGD = 8;
    static constexpr int S[NT][NC] =
        {{-1, 1}, {1, -1}};
    static constexpr int G[NT][NT] =
        {{1, 1}, {1, 1}};
    template <int J> SIMINF_MODEL_FN double rate(SIMINF_RATE_ARGS)
    {
        if constexpr (J == 0) return SISe_S_to_I(u, v, ld, gd, t);
        else return SISe_I_to_S(u, v, ld, gd, t);
    }
    SIMINF_MODEL_FN int pts(SIMINF_PTS_ARGS)
    {
        return SISe_pts(vn, u, v, ld, gd, node, t);
    }
};

struct ModelSISe3 {
    static constexpr int NC = 6, NT = 6, ND = 1, NLD = 4, NGD = 12;
    static constexpr int S[NT][NC] =
        {{-1, 1, 0, 0, 0, 0}, {1, -1, 0, 0, 0, 0}, {0, 0, -1, 1, 0, 0},
         {0, 0, 1, -1, 0, 0}, {0, 0, 0, 0, -1, 1}, {0, 0, 0, 0, 1, -1}};
    static constexpr int G[NT][NT] =
        {{1, 1, 0, 0, 0, 0}, {1, 1, 0, 0, 0, 0}, {0, 0, 1, 1, 0, 0},
         {0, 0, 1, 1, 0, 0}, {0, 0, 0, 0, 1, 1}, {0, 0, 0, 0, 1, 1}};
    template <int J> SIMINF_MODEL_FN double rate(SIMINF_RATE_ARGS)
    {
        if constexpr (J == 0) return SISe3_S_1_to_I_1(u, v, ld, gd, t);
        else if constexpr (J == 1) return SISe3_I_1_to_S_1(u, v, ld, gd, t);
        else if constexpr (J == 2) return SISe3_S_2_to_I_2(u, v, ld, gd, t);
        else if constexpr (J == 3) return SISe3_I_2_to_S_2(u, v, ld, gd, t);
        else if constexpr (J == 4) return SISe3_S_3_to_I_3(u, v, ld, gd, t);
        else return SISe3_I_3_to_S_3(u, v, ld, gd, t);
    }
    SIMINF_MODEL_FN int pts(SIMINF_PTS_ARGS)
    {
        return SISe3_pts(vn, u, v, ld, gd, node, t);
    }
};

/* The spatial models: SISe's and SISe3's rates, a post-time-step function
 * that reads the neighbours listed in ldata after end_t1..end_t4 (from
 * row NEIGHBOR_ROW).  NLD covers the four interval ends the rates'
 * context holds; the function receives the node's full ldata column. */
struct ModelSISe_sp {
    static constexpr int NC = 2, NT = 2, ND = 1, NLD = 4, NGD = 8;
    static constexpr bool NEIGHBORS = true;
    static constexpr int NEIGHBOR_ROW = 4;
    static constexpr int S[NT][NC] =
        {{-1, 1}, {1, -1}};
    static constexpr int G[NT][NT] =
        {{1, 1}, {1, 1}};
    template <int J> SIMINF_MODEL_FN double rate(SIMINF_RATE_ARGS)
    {
        if constexpr (J == 0) return SISe_S_to_I(u, v, ld, gd, t);
        else return SISe_I_to_S(u, v, ld, gd, t);
    }
    SIMINF_MODEL_FN int pts(SIMINF_PTS_ARGS)
    {
        return SISe_sp_pts(vn, u, v, ld, gd, node, t);
    }
};

struct ModelSISe3_sp {
    static constexpr int NC = 6, NT = 6, ND = 1, NLD = 4, NGD = 12;
    static constexpr bool NEIGHBORS = true;
    static constexpr int NEIGHBOR_ROW = 4;
    static constexpr int S[NT][NC] =
        {{-1, 1, 0, 0, 0, 0}, {1, -1, 0, 0, 0, 0}, {0, 0, -1, 1, 0, 0},
         {0, 0, 1, -1, 0, 0}, {0, 0, 0, 0, -1, 1}, {0, 0, 0, 0, 1, -1}};
    static constexpr int G[NT][NT] =
        {{1, 1, 0, 0, 0, 0}, {1, 1, 0, 0, 0, 0}, {0, 0, 1, 1, 0, 0},
         {0, 0, 1, 1, 0, 0}, {0, 0, 0, 0, 1, 1}, {0, 0, 0, 0, 1, 1}};
    template <int J> SIMINF_MODEL_FN double rate(SIMINF_RATE_ARGS)
    {
        if constexpr (J == 0) return SISe3_S_1_to_I_1(u, v, ld, gd, t);
        else if constexpr (J == 1) return SISe3_I_1_to_S_1(u, v, ld, gd, t);
        else if constexpr (J == 2) return SISe3_S_2_to_I_2(u, v, ld, gd, t);
        else if constexpr (J == 3) return SISe3_I_2_to_S_2(u, v, ld, gd, t);
        else if constexpr (J == 4) return SISe3_S_3_to_I_3(u, v, ld, gd, t);
        else return SISe3_I_3_to_S_3(u, v, ld, gd, t);
    }
    SIMINF_MODEL_FN int pts(SIMINF_PTS_ARGS)
    {
        return SISe3_sp_pts(vn, u, v, ld, gd, node, t);
    }
};

struct ModelSIRmparse {
    static constexpr int NC = 3, NT = 2, ND = 0, NLD = 0, NGD = 2;
    static constexpr bool PTS_NOOP = true;   /* pts never changes anything */
    static constexpr int S[NT][NC] =
        {{-1, 1, 0}, {0, -1, 1}};
    static constexpr int G[NT][NT] =
        {{1, 1}, {1, 1}};
    template <int J> SIMINF_MODEL_FN double rate(SIMINF_RATE_ARGS)
    {
        if constexpr (J == 0) return SIR_mparse_trFun1(u, v, ld, gd, t);
        else return SIR_mparse_trFun2(u, v, ld, gd, t);
    }
    SIMINF_MODEL_FN int pts(SIMINF_PTS_ARGS)
    {
        return siminf_pts_noop(vn, u, v, ld, gd, node, t);
    }
};

} /* namespace siminf */

/* The coupled kernels of a model are instantiated in a translation unit
 * of their own (builtin_*_coupled.hip), so that the library's units
 * compile in parallel: the record's unit declares the instantiation
 * extern, the other one defines it. */
#define SIMINF_COUPLED_INSTANCE(PREFIX, M)                                      \
    PREFIX template hipError_t siminf::launch_coupled<M>(                       \
        const siminf::KParams &, bool, bool, unsigned, unsigned, size_t, hipStream_t, \
        int *)
