// Plugin translation and device-program generation (host side, no HIP calls).
#pragma once

#include <string>
#include <vector>

#include "../../include/lensed.h"

namespace lsd {

// Turn an object name into a C identifier suffix ("sersic-old" -> "sersic_old").
std::string mangle(const std::string& name);

// Translate one plugin written in the OpenCL-C object dialect into C++ that
// compiles against kernels/prelude.h (inside namespace lcl).  `mangled` is
// the identifier suffix used for the object's symbols.
std::string translate_object(const std::string& source, const std::string& mangled);

// Wrap a translated object in the naming macros of the main program:
// type_<m>, parlst_<m>, data_<m>, deflection_<m>, brightness_<m>,
// foreground_<m>, set_<m>.
std::string wrap_object(const std::string& translated, const std::string& mangled,
                        const std::string& name);

// Wrap a translated object so that its metadata land in named device
// globals of a code object (lensed_meta_type, lensed_meta_size,
// lensed_meta_params).
std::string wrap_object_meta(const std::string& translated, const std::string& mangled,
                             const std::string& name);

// Locate and read an object file: the directories in LENSED_OBJECTS_PATH
// (colon separated) first, then the package's own objects/ directory.
// Returns false if no file is found.
bool load_object_source(const std::string& name, std::string* source, std::string* path);

// Directory that holds prelude.h and the kernel templates.
std::string kernel_dir();

// Read a whole text file; false if it cannot be opened.
bool read_text(const std::string& path, std::string* out);

// One object of a model, as the generator needs it.
struct ModelObject {
    std::string name;      // plugin name
    int type;              // OBJ_LENS / OBJ_SOURCE / OBJ_FOREGROUND
    size_t size;           // data size in 4-byte words (rounded up to 16 bytes)
    size_t npars;
    std::vector<int> ptype; // parameter types
    std::vector<int> ipp;   // image-plane prior flags
    bool table = false;     // EPL deflection with the tabulated angular factor
};

// Whether a lens plugin can take the tabulated EPL angular factor: the table
// form reads the lens's data block by field name, so its data block must
// declare float2 c (centre), mat22 fwd (image -> elliptical frame, scaled),
// mat22 bwd (back, with the normalisation), float coef[10] (the series'
// coefficients) and float hexp ((1 - t)/2); the caller also needs its
// parameters t and q.  With a mat22 shear field the shear is folded in too.
bool object_tabulates(const std::string& name);
bool object_table_shear(const std::string& name);

// Generated code for one model (ordered object list): compute_<k>() and
// set_params_<k>().  Offsets are in 4-byte words.
std::string compute_function(const std::vector<ModelObject>& objs, int k);
// whether the model's pixels can be culled before the render: a tabulated
// EPL lens and one source with a frame and a cutoff, nothing else
bool model_culls(const std::vector<ModelObject>& objs);
std::string set_params_function(const std::vector<ModelObject>& objs, int k);

// Total data words of a model.
size_t model_data_words(const std::vector<ModelObject>& objs);

// Prelude, translated objects and the generated compute_<k>/set_params_<k>
// for a list of models: compiles for the device and with a host compiler.
std::string model_source(const std::vector<std::vector<ModelObject>>& models, std::string* err);

// Assemble the whole device program (prelude, objects, model code, kernels)
// for a list of models; empty string and *err on failure.
// psfs: the (width, height, separable) PSF shapes that get compile-time
// specialised convolution code; any other shape runs the generic path.
struct PsfShape {
    int w, h, sep;
    bool operator==(const PsfShape& o) const { return w == o.w && h == o.h && sep == o.sep; }
};
std::string assemble_program(const std::vector<std::vector<ModelObject>>& models, std::string* err,
                             const std::vector<PsfShape>& psfs = {});

}  // namespace lsd
