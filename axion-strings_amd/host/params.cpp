// params.cpp -- parameter-file reader (reference src/common/
// read_parameters.cpp:30-238).  Same tag table, value types and error
// messages; the added tags are optional.
#include "params.hpp"

#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

namespace {

enum class Type { FLOAT, INT, UINT, STRING };

struct Tag {
    const char *name;
    Type type;
    void *addr;
    size_t cap;       // STRING capacity
    bool required;
    bool seen;
};

std::vector<Tag> tag_table(parameters *p)
{
    // order and types of read_parameters.cpp:36-169
    return {
        {"lambda", Type::FLOAT, &p->lambda, 0, true, false},
        {"enable_PRS", Type::INT, &p->enable_PRS, 0, true, false},
        {"T_initial", Type::FLOAT, &p->T_initial, 0, true, false},
        {"NDIMS", Type::INT, &p->NDIMS, 0, true, false},
        {"N", Type::INT, &p->N, 0, true, false},
        {"space_step", Type::FLOAT, &p->space_step, 0, true, false},
        {"time_step", Type::FLOAT, &p->time_step, 0, true, false},
        {"stencil_setting", Type::INT, &p->stencil_setting, 0, true, false},
        {"seed", Type::UINT, &p->seed, 0, true, false},
        {"write_output_file", Type::INT, &p->write_output_file, 0, true, false},
        {"output_file_path", Type::STRING, p->output_file_path, AXS_MAX_LEN, true, false},
        {"save_snapshots", Type::INT, &p->save_snapshots, 0, true, false},
        {"n_snapshots", Type::INT, &p->n_snapshots, 0, true, false},
        {"output_directory", Type::STRING, p->output_directory, AXS_MAX_LEN, true, false},
        {"save_fields", Type::INT, &p->save_fields, 0, true, false},
        {"save_strings", Type::INT, &p->save_strings, 0, true, false},
        {"save_pk", Type::INT, &p->save_pk, 0, true, false},
        {"sample_time_series", Type::INT, &p->sample_time_series, 0, true, false},
        {"n_samples", Type::INT, &p->n_samples, 0, true, false},
        {"ts_output_path", Type::STRING, p->ts_output_path, AXS_MAX_LEN, true, false},
        {"sample_strings", Type::INT, &p->sample_strings, 0, true, false},
        {"sample_background", Type::INT, &p->sample_background, 0, true, false},
        {"thr", Type::INT, &p->thr, 0, true, false},
        {"enable_amr", Type::INT, &p->enable_amr, 0, true, false},
        {"init_from_snapshot", Type::INT, &p->init_from_snapshot, 0, true, false},
        {"tau_initial", Type::FLOAT, &p->tau_initial, 0, true, false},
        {"refinement_threshold", Type::FLOAT, &p->refinement_threshold, 0, true, false},
        // optional additions
        {"n_steps", Type::INT, &p->n_steps, 0, false, false},
        {"precision", Type::STRING, p->precision, sizeof p->precision, false, false},
        {"init_mode", Type::STRING, p->init_mode, sizeof p->init_mode, false, false},
        {"init_on_device", Type::INT, &p->init_on_device, 0, false, false},
        {"sample_energy", Type::INT, &p->sample_energy, 0, false, false},
        {"pk_binning", Type::STRING, p->pk_binning, sizeof p->pk_binning, false, false},
        {"pk_nbins", Type::INT, &p->pk_nbins, 0, false, false},
        {"save_full_fields", Type::INT, &p->save_full_fields, 0, false, false},
        {"device", Type::INT, &p->device, 0, false, false},
        {"checkpoint_path", Type::STRING, p->checkpoint_path, AXS_MAX_LEN, false, false},
        {"restart_path", Type::STRING, p->restart_path, AXS_MAX_LEN, false, false},
        {"eom", Type::STRING, p->eom, sizeof p->eom, false, false},
        {"laplacian", Type::STRING, p->laplacian, sizeof p->laplacian, false, false},
        {"debug_nan", Type::INT, &p->debug_nan, 0, false, false},
    };
}

}  // namespace

void parameters_set_defaults(parameters *p)
{
    p->n_steps = 0;
    snprintf(p->precision, sizeof p->precision, "fp64");
    snprintf(p->init_mode, sizeof p->init_mode, "thermal");
    p->init_on_device = 0;
    p->sample_energy = 0;
    snprintf(p->pk_binning, sizeof p->pk_binning, "linear");
    p->pk_nbins = 64;
    p->save_full_fields = 0;
    p->device = 0;
    p->checkpoint_path[0] = 0;
    p->restart_path[0] = 0;
    snprintf(p->eom, sizeof p->eom, "field_rescaled");
    snprintf(p->laplacian, sizeof p->laplacian, "stencil");
    p->debug_nan = 0;
}

int read_parameter_file(const char *fname, parameters *p)
{
    memset(p, 0, sizeof *p);
    parameters_set_defaults(p);
    std::vector<Tag> tags = tag_table(p);

    FILE *fd = fopen(fname, "r");
    if (!fd) {
        printf("Error: cannot open parameter file '%s'.\n", fname);
        return 1;
    }
    char line[200], buf1[200], buf2[200], buf3[200];
    // fgets with the reference's 200-character line (MAX_LINE_WIDTH)
    while (fgets(line, sizeof line, fd) != NULL) {
        if (sscanf(line, "%199s%199s%199s", buf1, buf2, buf3) < 2 || buf1[0] == '%') continue;
        Tag *t = nullptr;
        for (Tag &x : tags)
            if (!x.seen && strcmp(buf1, x.name) == 0) {
                t = &x;
                break;
            }
        if (!t) {
            printf("Error in file %s:   Tag '%s' not allowed or multiple defined.\n", fname,
                   buf1);                                            // :222
            fclose(fd);
            return 1;
        }
        t->seen = true;
        switch (t->type) {
        case Type::FLOAT: *(float *)t->addr = (float)atof(buf2); break;
        case Type::INT: *(int *)t->addr = atoi(buf2); break;
        case Type::UINT: *(unsigned int *)t->addr = (unsigned int)atoi(buf2); break;
        case Type::STRING:
            if (strlen(buf2) >= t->cap) {
                printf("Error in file %s:   value of tag '%s' is too long.\n", fname, buf1);
                fclose(fd);
                return 1;
            }
            strcpy((char *)t->addr, buf2);
            break;
        }
    }
    fclose(fd);
    for (const Tag &x : tags)
        if (x.required && !x.seen) {
            printf("Error. I miss a value for tag '%s' in parameter file '%s'.\n", x.name,
                   fname);                                           // :233
            return 1;
        }
    return 0;
}
