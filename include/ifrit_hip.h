/*
 * C ABI of the MI355X tile rasterizer.
 *
 * The first block keeps the symbols, signatures and argument meaning of the
 * reference's exported C ABI, so that callers written against it (the C#
 * P/Invoke bindings in bindings/csharp/IfritLib/Native/NativeMethods*.cs and
 * any C caller) bind unchanged.  Handles the reference passes as C++ object
 * pointers (TileRasterRendererWrapper*, FrameBuffer*, ImageF32*, ...) are
 * opaque pointers here.  As in the reference, these functions return void
 * and do not throw; errors that the reference would abort on are recorded
 * and can be read with iftrGetLastError.
 *
 * The second block adds what the GPU backend needs on top: sample counts,
 * textures, built-in GPU shader programs, visibility-buffer readback and
 * frame statistics.
 */
#ifndef IFRIT_HIP_H
#define IFRIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct IfritRendererHandle IfritRendererHandle;       /* TileRasterRendererWrapper* */
typedef struct IfritFrameBuffer IfritFrameBuffer;             /* FrameBuffer* */
typedef struct IfritImageF32 IfritImageF32;                   /* Core::Data::ImageF32* */
typedef struct IfritVertexBuffer IfritVertexBuffer;           /* VertexBuffer* */
typedef struct IfritVaryingDescriptor IfritVaryingDescriptor; /* VaryingDescriptor* */

typedef struct {
    float x, y, z, w;
} IfritVector4f;

/* include/ifrit/softgraphics/engine/base/Shaders.h:116-118 */
typedef void (*IfritVertexShaderFunctionalPtr)(const void* const* input, IfritVector4f* outPos,
                                               IfritVector4f* const* outVaryings);
typedef void (*IfritFragmentShaderFunctionalPtr)(const void* varyings, void* colorOutput, float* fragmentDepth);

/* include/ifrit/softgraphics/engine/base/Structures.h:90-97 */
typedef struct {
    uint8_t blendEnable;
    int32_t srcColorBlendFactor;
    int32_t dstColorBlendFactor;
    int32_t srcAlphaBlendFactor;
    int32_t dstAlphaBlendFactor;
} IfritColorAttachmentBlendState;

/* IfritCompareOp, include/ifrit/softgraphics/engine/base/Constants.h:77-87 */
enum {
    IF_COMPARE_OP_NEVER = 0,
    IF_COMPARE_OP_LESS = 1,
    IF_COMPARE_OP_EQUAL = 2,
    IF_COMPARE_OP_LESS_OR_EQUAL = 3,
    IF_COMPARE_OP_GREATER = 4,
    IF_COMPARE_OP_NOT_EQUAL = 5,
    IF_COMPARE_OP_GREATER_OR_EQUAL = 6,
    IF_COMPARE_OP_ALWAYS = 7
};

/* TypeDescriptorEnum, include/ifrit/softgraphics/engine/base/TypeDescriptor.h:25-39 */
enum {
    IFTP_UNDEFINED = 0,
    IFTP_FLOAT1,
    IFTP_FLOAT2,
    IFTP_FLOAT3,
    IFTP_FLOAT4,
    IFTP_INT1,
    IFTP_INT2,
    IFTP_INT3,
    IFTP_INT4,
    IFTP_UINT1,
    IFTP_UINT2,
    IFTP_UINT3,
    IFTP_UINT4
};

/* ---- Renderer: include/ifrit/softgraphics/engine/export/TileRasterRendererExport.h:33-76 ---- */
IfritRendererHandle* iftrCreateInstance(void);                                      /* EXH:33 */
void iftrDestroyInstance(IfritRendererHandle* h);                                   /* EXH:35 */
void iftrBindFrameBuffer(IfritRendererHandle* h, IfritFrameBuffer* frameBuffer);    /* EXH:36-38 */
void iftrBindVertexBuffer(IfritRendererHandle* h, const IfritVertexBuffer* vb);     /* EXH:39-41 */
void iftrBindIndexBuffer(IfritRendererHandle* h, void* indexBuffer);                /* EXH:42-43, an ifbufCreateBuffer handle */
void iftrBindVertexShaderFunc(IfritRendererHandle* h, IfritVertexShaderFunctionalPtr func,
                              IfritVaryingDescriptor* vsOutDescriptors);             /* EXH:44-47 */
void iftrBindFragmentShaderFunc(IfritRendererHandle* h, IfritFragmentShaderFunctionalPtr func); /* EXH:48-50 */
void iftrSetBlendFunc(IfritRendererHandle* h, IfritColorAttachmentBlendState* state); /* EXH:51-53 */
void iftrSetDepthFunc(IfritRendererHandle* h, int32_t state);                       /* EXH:54-56 */
void iftrOptsetForceDeterministic(IfritRendererHandle* h, int opt);                 /* EXH:57-58 */
void iftrOptSetDepthTestEnable(IfritRendererHandle* h, int opt);                    /* EXH:59-60 */
/* The same entry under the name the reference's C# binding imports
   (bindings/csharp/Ifrit-v2-csharp/IfritLib/Native/NativeMethods.Engine.cs:103, EntryPoint
   "iftrOptsetDepthTestEnable"), so that binding resolves against this library unchanged. */
void iftrOptsetDepthTestEnable(IfritRendererHandle* h, int opt);
void iftrDrawLegacy(IfritRendererHandle* h, int numVertices, int clearFramebuffer); /* EXH:61-63 */
void iftrClear(IfritRendererHandle* h);                                             /* EXH:64-65 */
void iftrInit(IfritRendererHandle* h);                                              /* EXH:66-67 */
void iftrTest(void (*p)(int));                                                      /* EXH:68-69 */
void iftrBindVertexShader(IfritRendererHandle* h, void* pVertexShader);             /* EXH:72-73 */
void iftrBindFragmentShader(IfritRendererHandle* h, void* pFragmentShader);         /* EXH:74-75 */

/* ---- Frame buffers, varyings, vertex buffers: include/ifrit/softgraphics/engine/export/EngineCommonExport.h:33-70 ---- */
IfritFrameBuffer* iftrCreateFrameBuffer(void);
void iftrDestroyFrameBuffer(IfritFrameBuffer* p);
void iftrSetFrameBufferColorAttachmentFP32(IfritFrameBuffer* p, IfritImageF32* const* pImage, size_t nums);
void iftrSetFrameBufferDepthAttachmentFP32(IfritFrameBuffer* p, IfritImageF32* pImage);
IfritVaryingDescriptor* iftrCreateVaryingDescriptor(void);
void iftrDestroyVaryingDescriptor(IfritVaryingDescriptor* p);
void iftrWriteVaryingDescriptor(IfritVaryingDescriptor* p, const int* pDesc, size_t num);
IfritVertexBuffer* iftrCreateVertexBuffer(void);
void iftrDestroyVertexBuffer(IfritVertexBuffer* p);
void iftrSetVertexBufferLayout(IfritVertexBuffer* p, const int* pDesc, size_t num);
void iftrSetVertexBufferSize(IfritVertexBuffer* p, size_t num);
void iftrAllocateVertexBuffer(IfritVertexBuffer* p);
void iftrSetVertexBufferValueFloat4(IfritVertexBuffer* p, int index, int attr, void* value);

/* ---- Images: include/ifrit/softgraphics/core/export/CommonExport.h:26-31 ---- */
IfritImageF32* ifcrCreateImageFP32(size_t width, size_t height, size_t channel);
void ifcrDestroyImageFP32(IfritImageF32* p);
float* ifcrGetImageRawDataFP32(IfritImageF32* p);

/* ---- Buffers: include/ifrit/softgraphics/engine/export/BufferManagerExport.h:25-35 ---- */
void* ifbufCreateBufferManager(void);
void ifbufDestroyBufferManager(void* p);
void* ifbufCreateBuffer(void* pManager, size_t bufSize);
void ifbufDestroyBuffer(void* p);
void ifbufBufferData(void* pBuffer, const void* pData, size_t offset, size_t size);

/* ======================= MI355X backend additions ======================= */

/* Built-in GPU programs (host function pointers cannot run on the GPU). */
enum {
    IFHIP_VS_TRANSFORM = 0,       /* clip = mvp * (attribute 0 xyz, 1) */
    IFHIP_VS_CLIP_PASSTHROUGH = 1 /* clip = attribute 0 (xyzw), for pre-transformed vertices */
};
enum {
    IFHIP_FS_FLAT = 0,        /* colour = bytes of a hash of the primitive id */
    IFHIP_FS_BLINN_PHONG = 1, /* textured Blinn-Phong, directional and point lights */
    IFHIP_FS_DEPTH_ONLY = 2,  /* visibility and depth only */
    IFHIP_FS_GGX_LAYERED = 3, /* GGX over three layers: texture 0 albedo, 1 tangent-space normal,
                                 2 roughness (red); needs a tangent in the vertex record */
    IFHIP_FS_NEURAL = 4,      /* the neural material trained by iftrNeuralTrain (same inputs as
                                 IFHIP_FS_GGX_LAYERED; light 0 only) */
    IFHIP_FS_FLAT_ALPHA = 6,  /* IFHIP_FS_FLAT with alpha = (64 + (hash byte 3 & 127)) / 255 */
    IFHIP_FS_NORMAL_LAMBERT = 7, /* untextured: d = max(0, n.l / |n|) with n the interpolated normal and
                                   l = lightPos[0].xyz as given; rgb = ambient + lightColor[0].rgb * d
                                   + d^shininess, alpha 1 (the shader of the reference CPU path's
                                   timing harness) */
    IFHIP_FS_CUBE_MAP = 8       /* colour = the cube map (a six-layer texture, the one trisPerMaterial
                                   selects) along the interpolated normal at level of detail `lod`
                                   (textureCubeLod, ShaderBuiltinCuda.cuh:184-231, 379-392) */
};
/* IfritBlendFactor, include/ifrit/softgraphics/engine/base/Constants.h:64-76.  Blending
   (iftrSetBlendFunc) supports ZERO, ONE, SRC_ALPHA, ONE_MINUS_SRC_ALPHA, DST_ALPHA and
   ONE_MINUS_DST_ALPHA, with the reference's encoding (TileRasterRenderer.cpp:212-333), at one
   sample per pixel for a single target; fragments are blended in primitive order. */
enum {
    IF_BLEND_FACTOR_ZERO = 0,
    IF_BLEND_FACTOR_ONE = 1,
    IF_BLEND_FACTOR_SRC_ALPHA = 6,
    IF_BLEND_FACTOR_ONE_MINUS_SRC_ALPHA = 7,
    IF_BLEND_FACTOR_DST_ALPHA = 8,
    IF_BLEND_FACTOR_ONE_MINUS_DST_ALPHA = 9
};
enum { IFHIP_SAMPLER_BILINEAR = 0, IFHIP_SAMPLER_TRILINEAR = 1, IFHIP_SAMPLER_ANISOTROPIC = 2 };

typedef struct {
    int32_t kind;     /* IFHIP_VS_TRANSFORM */
    float mvp[16];    /* row-major, applied as mvp * (x, y, z, 1) */
} IfritGpuVertexProgramDesc;

typedef struct {
    int32_t kind;               /* IFHIP_FS_* */
    int32_t offWorld, offNormal, offUv; /* float offsets in the vertex record */
    int32_t trisPerMaterial;    /* texture index = primitive / trisPerMaterial (0: texture 0) */
    int32_t sampler;            /* IFHIP_SAMPLER_* */
    int32_t maxAniso;         /* anisotropic taps, clamped to [1, 16] */
    float camera[3];
    float ambient, kd, ks, shininess;
    int32_t numLights;          /* <= 4 */
    float lightPos[4][4];       /* w = 0: direction towards the light, w = 1: point light */
    float lightColor[4][4];     /* rgb, attenuation k of 1 / (1 + k d^2) */
    int32_t offTangent;         /* float offset of the tangent (xyz), IFHIP_FS_GGX_LAYERED */
    int32_t precision;          /* IFHIP_FS_NEURAL: 0 BF16 MFMA, 1 FP8 (E4M3) MFMA for the hidden layer,
                                   BF16 for the input and output layers */
    int32_t samplerId;          /* 0: `sampler` / `maxAniso` above with repeat addressing and
                                   bilinear texels; else an id from iftrCreateSampler (Blinn-Phong and
                                   cube-map programs) */
    float lod;                  /* IFHIP_FS_CUBE_MAP: level of detail */
    int32_t depthWrite;         /* 1: the program writes depth' = depth * depthScale + depthBias through
                                   the fragment shader's depth out-parameter (Shaders.h:64).  As in the
                                   CPU path (TileRasterWorker.cpp:2272-2384) the depth test uses the
                                   interpolated depth and the sample then stores depth'; the draw folds
                                   fragments in primitive order.  The tagging mode's shading pass
                                   discards written depth (TileRasterWorker.cpp:2015-2068), so it is
                                   ignored there. */
    float depthScale, depthBias;
} IfritGpuFragmentProgramDesc;

void* iftrCreateGpuVertexProgram(const IfritGpuVertexProgramDesc* desc);
void* iftrCreateGpuFragmentProgram(const IfritGpuFragmentProgramDesc* desc);
void iftrDestroyGpuProgram(void* program);

/* IfritCullMode, include/ifrit/softgraphics/engine/base/Constants.h:29-35 */
enum {
    IF_CULL_MODE_NONE = 0,
    IF_CULL_MODE_FRONT = 1,
    IF_CULL_MODE_BACK = 2,
    IF_CULL_MODE_FRONT_AND_BACK = 3
};
/* SetCullMode of the CUDA renderer (TileRasterRendererCuda.h:109; behaviour of
   TileRasterCoreInvocationCuda.cu:1433-1462).  BACK (the default) is the CPU path's back-face cull.
   FRONT sets every triangle up with its first and third vertex exchanged, so only back faces
   survive; NONE does so for the triangles (and clip fans) that BACK would cull; FRONT_AND_BACK
   draws nothing.  Exchanged triangles shade with their own vertex attributes. */
void iftrSetCullMode(IfritRendererHandle* h, int mode);

/* Whose RCPPS the triangle setup reproduces.  The CPU path divides by w and by the signed area with
   _mm256_rcp_ps / _mm_rcp_ss and one Newton-Raphson step (TileRasterWorker.cpp:508-533,
   FastUtilSimd.h:27-33, 87-97); the approximation is implementation-defined, so the setup values,
   and with them edge-boundary pixels, depend on the CPU the reference runs on.  INTEL (the
   default) is the table of an Intel Xeon (top 11 mantissa bits), AMD that of an AMD EPYC (top 12),
   HOST picks the one of the CPU this process runs on (an error on other vendors).  No counterpart
   in the reference's interface, whose CPU answers the question by itself. */
enum { IFHIP_RCP_INTEL = 0, IFHIP_RCP_AMD = 1, IFHIP_RCP_HOST = 2 };
void iftrSetSetupReciprocal(IfritRendererHandle* h, int model);

/* Device buffers of the CUDA renderer (CreateBuffer / copyHostBufferToBuffer,
   TileRasterRendererCuda.h:98-99): numbered slots that mesh programs read.  CreateBuffer
   (re)allocates a slot of `size` bytes (contents undefined); copyHostBufferToBuffer writes
   `size` bytes at its start.  Returns 0, or -1 with an error. */
int iftrCreateBuffer(IfritRendererHandle* h, int slot, int size);
void iftrCopyHostBufferToBuffer(IfritRendererHandle* h, const void* src, int slot, int size);

/* Built-in mesh and task programs for DrawMeshTasks (bindMeshShader / bindTaskShader /
   DrawMeshTasks, TileRasterRendererCuda.h:88-89, 115; the pipeline of
   TileRasterCoreInvocationCuda.cu:355-500).  The reference runs user device objects; here a mesh
   work group runs the built-in meshlet program of the reference's meshlet demo
   (demo/shader/MeshletDemoShaders.cu): meshlet m owns vertices vo[m] .. vo[m + 1] - 1 of the vertex
   slot (records of recordFloats floats, the position xyz first) and indices io[m] .. io[m + 1] - 1 of
   the index slot, local to the meshlet's vertices, where vo and io are the int32 offset slots.  It
   emits each vertex as its record followed by mvp * (x, y, z, 1) (the transform program's summation
   order), so the fragment program reads the record at its usual offsets, and its indices rebased
   onto the emitted vertices.  The task program of that demo launches one
   mesh work group per task work group with the task's id as payload, i.e. meshlet m = work group m.
   The emitted meshlets are concatenated in work-group order (the reference's compaction takes
   them in atomic order) and drawn as a triangle list by the bound pipeline state: the primitive
   id of a triangle is its position in that list.  View batches (iftrSetViews) are refused. */
enum { IFHIP_MS_MESHLET = 1 };
enum { IFHIP_TS_ONE_MESH_GROUP = 1 };
typedef struct {
    int32_t kind;                /* IFHIP_MS_MESHLET */
    float mvp[16];               /* row-major */
    int32_t vertexSlot, indexSlot, vertexOffsetSlot, indexOffsetSlot;  /* the demo: 0, 1, 2, 3 */
    int32_t recordFloats;        /* floats per vertex record (>= 3) */
} IfritGpuMeshProgramDesc;
typedef struct {
    int32_t kind;                /* IFHIP_TS_ONE_MESH_GROUP */
} IfritGpuTaskProgramDesc;
void* iftrCreateGpuMeshProgram(const IfritGpuMeshProgramDesc* desc);
void* iftrCreateGpuTaskProgram(const IfritGpuTaskProgramDesc* desc);
void iftrBindMeshShader(IfritRendererHandle* h, void* program);   /* NULL unbinds */
void iftrBindTaskShader(IfritRendererHandle* h, void* program);   /* NULL unbinds */
/* DrawMeshTasks: numWorkGroups task (or, without a task program, mesh) work groups from
   firstWorkGroup (the reference accepts only 0; any start is taken here).  clear: 1 starts from
   the clear values, 0 draws over the current frame (the CUDA renderer's DrawMeshTasks does not
   clear; it is preceded by clear()). */
void iftrDrawMeshTasks(IfritRendererHandle* h, int numWorkGroups, int firstWorkGroup, int clear);

/* Built-in geometry program (bindGeometryShader, TileRasterRendererCuda.h:87; the geometry stage
   of TileRasterCoreInvocationCuda.cu:500-560).  The reference runs a user device object per
   triangle; here IFHIP_GS_CLIP_OFFSET is the reference demo's geometry shader
   (demo/shader/DefaultDemoShaders.cu:82-99): each triangle is emitted once, its three clip positions
   moved by `offset` (the demo: x + 0.03), its vertex records passed on.  With a geometry program
   bound, a draw emits the triangles in order (the order of the reference's deterministic sort keys,
   :539-542) as unshared vertices and draws them with the bound fragment program and state; the
   primitive id is the triangle's index.  The bound vertex program runs first (a GPU program; host
   vertex functions, view batches and iftrDrawNoAlloc are refused).  NULL unbinds. */
enum { IFHIP_GS_CLIP_OFFSET = 1 };
typedef struct {
    int32_t kind;       /* IFHIP_GS_CLIP_OFFSET */
    float offset[4];    /* added to every emitted clip position */
} IfritGpuGeometryProgramDesc;
void* iftrCreateGpuGeometryProgram(const IfritGpuGeometryProgramDesc* desc);
void iftrBindGeometryShader(IfritRendererHandle* h, void* program);

/* IfritPolygonMode, include/ifrit/softgraphics/engine/base/Constants.h:23-28 */
enum { IF_POLYGON_MODE_FILL = 0, IF_POLYGON_MODE_LINE = 1, IF_POLYGON_MODE_POINT = 2 };
/* setRasterizerPolygonMode of the CUDA renderer (TileRasterRendererCuda.h:106), after its point and
   line pipelines (TileRasterCoreInvocationCuda.cu:2780-3335), at 1 sample per pixel on a single
   target without blending.  POINT draws every index i of whole triangles as a point: clip position divided by w
   (points with w < 1e-8 dropped), mapped to [0, 1] and culled outside it, at pixel
   (int(x (W - 1)), int(y (H - 1))) with depth z / w; its primitive id is i.  LINE draws the three
   edges (v0 v1, v1 v2, v2 v0) of every triangle: an end behind w = 0 is cut at w = 1, and a line is
   culled when its first end leaves [0, 1]^2 or its second end has x < 0 or y < 0 (the spec's test);
   its pixels are the spec's Bresenham walk, each with the depth interpolated along the line's major
   axis; its primitive id is 3 * triangle + edge.  Per pixel the nearest candidate below 1e9 wins
   (equal depths: the lowest id) when it is below the stored depth; the depth function is not used,
   as in the spec.  The winner is shaded with the bound program from the point's vertex, or from the
   line's two ends perspective-correctly; textures are sampled without screen-space derivatives.
   Depth -0 is stored as +0, and pixels the spec would write outside the frame are dropped. */
void iftrSetRasterizerPolygonMode(IfritRendererHandle* h, int mode);

/* setClearValues of the CUDA renderer (TileRasterRendererCuda.h:111): the colour and depth a draw
   with clear starts from.  The first colour applies to the (one) colour attachment; samples no
   triangle covers resolve with it.  The defaults are the CPU path's: colour 0 and depth
   0x7f7f7f7f (TileRasterWorker.cpp:40, 113-118).  iftrClear keeps the CPU path's values. */
void iftrSetClearValues(IfritRendererHandle* h, const IfritVector4f* colors, size_t numColors, float depth);

/* SetScissors / setScissorTestEnable of the CUDA renderer (TileRasterRendererCuda.h:101-102):
   up to 8 rectangles (x, y, width, height) in pixels of the render target (or of each view of a
   batch); with the test on, a pixel (px, py) passes when it lies in one of them, x <= px <= x +
   width and y <= py <= y + height (TileRasterCoreInvocationCuda.cu:1699-1711).  A draw leaves
   the pixels that fail as they are: visibility, depth and colour, its clear included (the CUDA
   clear kernels are scissored too, :3336-3424).  The test is off by default; on with no
   rectangle, every pixel passes. */
void iftrSetScissors(IfritRendererHandle* h, const IfritVector4f* rects, size_t numRects);
void iftrSetScissorTestEnable(IfritRendererHandle* h, int enable);

/* setMsaaSamples of the CUDA renderer (TileRasterRendererCuda.h:105): 1, 4 or 8. */
void iftrSetMsaaSamples(IfritRendererHandle* h, int samples);
/* Attribute-0 stride used by the built-in vertex program when the bound vertex buffer is
   given as raw floats (iftrBindVertexDataFloat). */
void iftrBindVertexDataFloat(IfritRendererHandle* h, const float* data, int strideFloats, int numVertices);
void iftrBindIndexData(IfritRendererHandle* h, const int32_t* data, int numIndices);

/* Textures (CreateTexture / copyHostBufferToImage / generateMipmap of the CUDA renderer,
   TileRasterRendererCuda.h:90-96), RGBA8 with a full mip chain, power-of-two sizes.  A texture
   may hold several layers (IfritImageCreateInfo::arrayLayers, Structures.h:62-69): level l holds
   max(1, w >> l) x max(1, h >> l) texels per layer, its layers one after another; a cube map is
   six layers, +x -x +y -y +z -z. */
int iftrCreateTexture(IfritRendererHandle* h, int width, int height, int levels);
int iftrCreateTextureLayers(IfritRendererHandle* h, int width, int height, int levels, int layers);
void iftrCopyHostBufferToImageLayer(IfritRendererHandle* h, int texture, int layer, const void* rgba8Level0);

/* IfritFilter, IfritSamplerAddressMode, IfritBorderColor (Constants.h:36-53) and IfritSamplerT
   (Structures.h:52-61): CreateSampler of the CUDA renderer (TileRasterRendererCuda.h:92). */
enum { IF_FILTER_NEAREST = 0, IF_FILTER_LINEAR = 1 };
enum {
    IF_SAMPLER_ADDRESS_MODE_REPEAT = 0,
    IF_SAMPLER_ADDRESS_MODE_MIRRORED_REPEAT = 1,
    IF_SAMPLER_ADDRESS_MODE_CLAMP_TO_EDGE = 2,
    IF_SAMPLER_ADDRESS_MODE_CLAMP_TO_BORDER = 3,
    IF_SAMPLER_ADDRESS_MODE_MIRROR_CLAMP_TO_EDGE = 4
};
enum { IF_BORDER_COLOR_BLACK = 0, IF_BORDER_COLOR_WHITE = 1 };
typedef struct {
    int32_t filterMode;                         /* within a level: NEAREST texel or LINEAR (bilinear) */
    int32_t addressModeU, addressModeV, addressModeW;  /* W is kept, 2D and cube lookups use U and V */
    int32_t borderColor;                        /* CLAMP_TO_BORDER: opaque black or white */
    int32_t anisotropyEnable;                   /* 1: anisotropic taps (up to maxAnisotropy, at most 16) */
    float maxAnisotropy;
} IfritSamplerDesc;
/* Returns the sampler id (1, 2, ...) for IfritGpuFragmentProgramDesc::samplerId (-1 on error).  Texel centres
   are at (i + 0.5) / size; mirrored repeat reflects every other period and mirror-clamp-to-edge
   reflects once about 0 and then clamps (the Vulkan definitions); clamp-to-border gives border
   texels to the filter one by one. */
int iftrCreateSampler(IfritRendererHandle* h, const IfritSamplerDesc* desc);

/* blitImage (TileRasterRendererCuda.h:95; IfritImageBlit, Structures.h:70-83): a region of one level
   and layer into a region of another, scaled; destination texel (dst start + (x, y)) reads the
   source at (x / dst width * src width + src start, likewise y); IF_FILTER_LINEAR blends the four
   neighbours (right and lower clamped to the source level) and rounds half up, IF_FILTER_NEAREST
   takes the texel at the rounded coordinate. */
typedef struct {
    uint32_t mipLevel, baseArrayLayer;
} IfritImageSubresourceLayers;
typedef struct {
    uint32_t width, height, depth;
} IfritExtent3D;
typedef struct {
    IfritImageSubresourceLayers srcSubresource;
    IfritExtent3D srcExtentSt, srcExtentEd;
    IfritImageSubresourceLayers dstSubresource;
    IfritExtent3D dstExtentSt, dstExtentEd;
} IfritImageBlit;
void iftrBlitImage(IfritRendererHandle* h, int srcTexture, int dstTexture, const IfritImageBlit* region, int filter);
void iftrCopyHostBufferToImage(IfritRendererHandle* h, int texture, const void* rgba8Level0);
void iftrUploadTextureLevels(IfritRendererHandle* h, int texture, const void* rgba8AllLevels);
void iftrGenerateFbmTexture(IfritRendererHandle* h, int texture, uint32_t seed);
void iftrGenerateMipmap(IfritRendererHandle* h, int texture);
void iftrDownloadTexture(IfritRendererHandle* h, int texture, void* rgba8AllLevels);

/* Render target size without host attachments (device-resident frames). */
void iftrSetRenderTarget(IfritRendererHandle* h, int width, int height);

/* Frame outputs (device to host copies). */
void iftrGetVisibility(IfritRendererHandle* h, uint32_t* primIds, float* depth); /* W*H*S each */
void iftrGetCoverageClass(IfritRendererHandle* h, uint8_t* classes);            /* W*H, 1 sample */
void iftrGetColorRGBA8(IfritRendererHandle* h, uint8_t* rgba);                  /* W*H*4 */
/* Unquantised colour (W*H*4 floats): written by blending draws, or by every draw after
   iftrSetColorFloatOutput(h, 1). */
void iftrSetColorFloatOutput(IfritRendererHandle* h, int enable);
void iftrGetColorRGBA32F(IfritRendererHandle* h, float* rgba);

typedef struct {
    int64_t inputTriangles;
    int64_t assembledTriangles;
    int64_t clippedTriangles;
    int64_t binEntries;
    int64_t fragments;
    int64_t coveredSamples;
    uint32_t status;       /* overflow bits, 0 when the frame is complete */
    uint32_t retries;      /* frames re-run with larger buffers */
    float stageMs[8];      /* vertex (and counter reset), setup (clip, cull), bin, raster, raster merge,
                              shade, (unused), (unused) */
} IfritFrameStats;

void iftrGetFrameStats(IfritRendererHandle* h, IfritFrameStats* out);
/* Per-stage timing with device events (adds event records between stages). */
void iftrSetStageTiming(IfritRendererHandle* h, int enable);
/* Draw without waiting for completion, then wait; for timing loops. */
void iftrDrawAsync(IfritRendererHandle* h, int numVertices, int clearFramebuffer);
void iftrSynchronize(IfritRendererHandle* h);
/* NULL when no error has been recorded since the last call. */
const char* iftrGetLastError(IfritRendererHandle* h);

/* Neural material: an 8-channel FP16 latent texture with mips plus an MLP
   14 -> 64 -> 64 -> 3 (ReLU) on BF16 / FP8 MFMA, trained on fragments of the bound
   layered-GGX material (IFHIP_FS_GGX_LAYERED must be bound, with its three textures)
   rendered from random views of the bound geometry.  No counterpart in the reference. */
typedef struct {
    int32_t views;                    /* training views */
    int32_t viewWidth, viewHeight;    /* training view size (1 sample per pixel) */
    int32_t maxFragments;             /* data set capacity */
    int32_t steps, batch;             /* Adam steps, fragments per step (multiple of 256) */
    float lrLatent, lrMlp;            /* initial learning rates (cosine decay to 5 %) */
    uint32_t seed;
    int32_t latentSize;               /* power of two */
    float eyeDistMin, eyeDistMax;     /* camera distance from the origin */
    float fovY;                       /* radians; the largest field of view */
    float targetJitter;               /* look-at point uniform in [-j, j]^3 */
    float fovYMin;                    /* > 0: field of view log-uniform in [fovYMin, fovY] per view */
    int32_t singleTap;                /* 1: one trilinear latent tap at the isotropic footprint,
                                         0: the material's anisotropic taps */
} IfritNeuralTrainDesc;
typedef struct {
    int64_t fragments;                /* training fragments collected */
    float lossFirst, lossLast;        /* mean squared error per channel over the first / last step */
    float datasetMs, trainMs;
} IfritNeuralTrainReport;
void iftrNeuralTrain(IfritRendererHandle* h, const IfritNeuralTrainDesc* desc, IfritNeuralTrainReport* report);
/* Parameters as 5508 floats: W1[16][64], b1[64], W2[64][64], b2[64], W3[64][4], b3[4]. */
void iftrNeuralSetParams(IfritRendererHandle* h, const float* params);
void iftrNeuralGetParams(IfritRendererHandle* h, float* params);
/* Renders the current view with the layered-GGX material and with the neural material at the given
   precision (0 BF16, 1 FP8) and returns the PSNR (dB, RGB over covered pixels) of the neural frame
   against the material's.  Below minPsnrDb, IFHIP_FS_NEURAL at that precision falls back to the
   layered-GGX material until the next validation or training. */
float iftrNeuralValidate(IfritRendererHandle* h, int precision, float minPsnrDb, int numIndices);
/* iftrNeuralValidate with a per-pixel bound as well: the neural frame is compared with the
   material's over the covered pixels; below minPsnrDb, or with any pixel whose largest RGB error
   exceeds maxErrLsb RGBA8 steps (255: no per-pixel bound), IFHIP_FS_NEURAL at that precision falls
   back to the layered-GGX material until the next validation or training. */
typedef struct {
    float psnrDb;                     /* RGB over covered pixels */
    int32_t maxErrLsb;                /* largest per-pixel RGB error, RGBA8 steps */
    int32_t p999ErrLsb;               /* error that 99.9 % of the covered pixels stay within */
    int32_t fallback;                 /* 1: the neural program now draws the material */
    int64_t pixels;                   /* covered pixels compared */
} IfritNeuralValidation;
void iftrNeuralValidateEx(IfritRendererHandle* h, int precision, float minPsnrDb, int maxErrLsb, int numIndices,
                          IfritNeuralValidation* out);
/* The MLP alone on n inputs ([n][16] floats, 14 used) -> [n][4] outputs (3 used). */
void iftrNeuralEval(IfritRendererHandle* h, const float* in, float* out, int n, int precision);

/* Multi-view batches: every draw renders all views of the batch from the bound vertex and index
   data, each stage one launch over all views (the raster stage one persistent launch over the
   tiles of every view).  The bound GPU vertex program supplies the kind of transform and the
   bound fragment program the shading of views with shade = 1; views with shade = 0 produce
   visibility and FP32 depth only.  The MSAA sample count applies to every view.  No
   counterpart in the reference. */
typedef struct {
    int32_t width, height;            /* 1..16384 */
    float mvp[16];                    /* row-major, clip = mvp * (x, y, z, 1) */
    float camera[3];                  /* eye position for view-dependent shading */
    int32_t shade;                    /* 1: run the bound fragment program, 0: depth only */
} IfritViewDesc;
/* numViews = 0 returns to the single target of iftrBindFrameBuffer / iftrSetRenderTarget. */
void iftrSetViews(IfritRendererHandle* h, const IfritViewDesc* views, int numViews);
void iftrGetViewVisibility(IfritRendererHandle* h, int view, uint32_t* primIds, float* depth); /* W*H*S each */
void iftrGetViewColorRGBA8(IfritRendererHandle* h, int view, uint8_t* rgba);                  /* W*H*4 */
/* 1: each draw configuration's launches are captured once in a hipGraph, later draws replay it. */
void iftrSetGraphMode(IfritRendererHandle* h, int enable);

/* Launch parameters that change speed but never results.  The 16-pixel coverage tile is fixed:
   the reference's tile and block classification decides which depth expression a pixel uses. */
typedef struct {
    int32_t rasterPartEntries;        /* bin-list entries per raster work item: 256, 512 or 1024 */
    int32_t rasterGroupsPerCU;        /* persistent raster workgroups per compute unit, 1..8 */
    int32_t shadeTile;                /* shading tile side in pixels: 8, 16 or 32 */
    int32_t bigTriangleTiles;         /* triangles over more tiles are binned by a whole workgroup */
    uint32_t binCapacity;             /* bin-list entries reserved (grown on demand when exceeded) */
    int32_t rasterTileThreads;        /* threads per 16x16 tile at 1 sample per pixel: 256, or 64 (one
                                         wave per tile, more tiles in flight); other sample counts use 256 */
} IfritTuning;
void iftrSetTuning(IfritRendererHandle* h, const IfritTuning* tuning);
void iftrGetTuning(IfritRendererHandle* h, IfritTuning* tuning);
typedef struct {
    int32_t candidates;               /* configurations timed */
    int32_t identical;                /* 1: every configuration produced the same outputs */
    float defaultMs, bestMs;          /* mean draw time with the starting and the chosen tuning */
    uint32_t binDemand;               /* bin-list entries the draw needs */
    IfritTuning best;                 /* chosen and applied */
    IfritTuning tried[64];
    float ms[64];
} IfritAutotuneReport;
/* Coordinate search over the tuning parameters, each configuration timed over `frames` graph
   replays of the current draw (batch or single view); the bin capacity is then set to the
   measured demand plus 1/8. */
void iftrAutotune(IfritRendererHandle* h, int numIndices, int frames, IfritAutotuneReport* report);
/* Order-independent 64-bit digest of the last draw's depth, primitive IDs and colour. */
uint64_t iftrGetDigest(IfritRendererHandle* h);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
   hipGraph capture, tight serving loops, harnesses that forbid hipMalloc): the draw of
   iftrDrawLegacy for the bound single target (iftrSetRenderTarget or iftrBindFrameBuffer size,
   MSAA sample count, GPU vertex and fragment programs, textures, depth function), from the
   caller's device-resident vertices and indices into the caller's device outputs.  It calls no
   hipMalloc or hipFree: every intermediate buffer (clip positions, triangle records, bins, raster
   queue, counters, the per-sample record slots) is carved out of one static device array of
   IFHIP_NOALLOC_SCRATCH_BYTES per GPU, which the library holds from iftrInit until the process
   ends.  A draw whose buffers do not fit is refused with an error (iftrGetLastError) before any
   launch; 1920x1080 at 1 sample with 128K triangles, or 1024x1024 at 4 samples with 64K
   triangles, fit.  Every call on a GPU uses that GPU's array and keeps the last call's contents:
   calls that overlap on one GPU (two renderers, two threads) race.  The frame is always cleared;
   it blocks until the frame is done.  No blending, host vertex functions or view batches.
   Overflowing the bin or clip capacities (which a normal draw grows) is reported as an error. */
#define IFHIP_NOALLOC_SCRATCH_BYTES (48u << 20)
typedef struct {
    const float* vertices;      /* device: numVertices records of strideFloats floats */
    int32_t strideFloats;
    int32_t numVertices;
    const int32_t* indices;     /* device: numIndices int32 */
    int32_t numIndices;
    uint32_t* primIds;          /* device, W*H*S primitive IDs (0xffffffff: empty), or NULL */
    float* depth;               /* device, W*H*S FP32 depth */
    uint8_t* rgba8;             /* device, W*H*4 resolved colour, or NULL to skip shading */
} IfritNoAllocDraw;
void iftrDrawNoAlloc(IfritRendererHandle* h, const IfritNoAllocDraw* draw);

#ifdef __cplusplus
}
#endif

#endif /* IFRIT_HIP_H */
