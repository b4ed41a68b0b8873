# Builds the reference MonoAlg3D CPU path from its own sources, in place, into
# oracle/_ref/ (never committed).  Test infrastructure only: the reference binary
# and plugins are the checker for the oracle restatement and the "reference"
# CPU baseline; nothing in the product links them.
#
# Usage:  make -f oracle/ref.mk -j8            (from the repository root)
#
# Mirrors what the reference's build.sh does for the `simulator` target with no
# CUDA (build.sh:46-49,114: -O3 -fopenmp -DCELL_MODEL_REAL_DOUBLE), restricted to
# the plugins this path needs.  Only gcc is used; no reference file is copied.

REF ?= /root/reference
OUT ?= oracle/_ref
SRC := $(REF)/src

CC      ?= gcc
CFLAGS  := -O3 -fopenmp -std=gnu99 -fno-strict-aliasing -fPIC -w -DCELL_MODEL_REAL_DOUBLE
LDLIBS  := -ldl -lm -fopenmp

OBJDIR  := $(OUT)/obj
LIBDIR  := $(OUT)/shared_libs
BINDIR  := $(OUT)/bin

# Core static pieces (build.sh:171-186 STATIC_DEPS of MonoAlg3D).
CORE_SRCS := \
  3dparty/sds/sds.c config_helpers/config_helpers.c \
  utils/search.c utils/stop_watch.c utils/sort.c utils/file_utils.c utils/batch_utils.c \
  alg/grid/grid.c alg/grid/grid_refinement.c alg/grid/grid_derefinement.c \
  alg/cell/cell.c alg/cell/cell_derefinement.c alg/cell/cell_refinement.c \
  alg/grid_purkinje/grid_purkinje.c \
  monodomain/monodomain_solver.c ode_solver/ode_solver.c 3dparty/ini_parser/ini.c \
  config/config_common.c config/config_parser.c config/postprocessor_config.c \
  graph/graph.c graph/pqueue.c 3dparty/xml_parser/yxml.c 3dparty/tinyexpr/tinyexpr.c \
  3dparty/miniz/miniz.c vtk_utils/pvd_utils.c vtk_utils/data_utils.c \
  vtk_utils/vtk_polydata_grid.c vtk_utils/vtk_unstructured_grid.c ensight_utils/ensight_grid.c

CORE_OBJS := $(addprefix $(OBJDIR)/,$(CORE_SRCS:.c=.o))

# plugin name -> source list (relative to src/)
PLUG_default_stimuli                := stimuli_library/stimuli.c
PLUG_default_domains                := domains_library/domain.c domains_library/domain_helpers.c
PLUG_default_matrix_assembly        := matrix_assembly_library/matrix_assembly.c
PLUG_default_linear_system_solver   := linear_system_solver_library/linear_system_solver.c
PLUG_default_save_mesh              := save_mesh_library/save_mesh_helper.c save_mesh_library/save_mesh.c
PLUG_default_update_monodomain      := update_monodomain_library/update_monodomain.c
PLUG_ten_tusscher_2006              := models_library/ten_tusscher/ten_tusscher_2006_RS_CPU.c
PLUG_ohara_rudy_endo_2011           := models_library/Ohara_Rudy/Ohara_Rudy_2011.c
PLUG_mitchell_shaeffer_2003         := models_library/mitchell_shaeffer/mitchell_shaeffer_2003.c
PLUG_courtemanche_ramirez_nattel_1998 := models_library/courtemanche_ramirez_nattel_1998/courtemanche_ramirez_nattel_1998.c
PLUG_default_save_state             := save_state_library/save_state.c
PLUG_default_restore_state          := restore_state_library/restore_state.c

PLUGINS := default_stimuli default_domains default_matrix_assembly default_linear_system_solver \
           default_save_mesh default_update_monodomain ten_tusscher_2006 ohara_rudy_endo_2011 \
           mitchell_shaeffer_2003 courtemanche_ramirez_nattel_1998 default_save_state default_restore_state

PLUGIN_SOS := $(addprefix $(LIBDIR)/lib,$(addsuffix .so,$(PLUGINS)))

all: $(BINDIR)/MonoAlg3D $(PLUGIN_SOS) $(BINDIR)/ref_harness

# golden-vector harness (oracle/ref_harness.c) over the reference plugins
$(BINDIR)/ref_harness: oracle/ref_harness.c $(OUT)/libcore.a $(LIBDIR)/liblogger.so
	@mkdir -p $(BINDIR)
	$(CC) $(CFLAGS) -I$(SRC) -o $@ oracle/ref_harness.c $(OUT)/libcore.a -L$(LIBDIR) -llogger \
	    -Wl,-rpath,'$$ORIGIN/../shared_libs' $(LDLIBS)

$(OBJDIR)/%.o: $(SRC)/%.c
	@mkdir -p $(dir $@)
	$(CC) $(CFLAGS) -c $< -o $@

$(LIBDIR)/liblogger.so: $(OBJDIR)/logger/logger.o $(OBJDIR)/3dparty/sds/sds.o
	@mkdir -p $(LIBDIR)
	$(CC) -shared -o $@ $^ $(LDLIBS)

$(OUT)/libcore.a: $(CORE_OBJS)
	@mkdir -p $(OUT)
	rm -f $@ && ar rcs $@ $^

$(BINDIR)/MonoAlg3D: $(OBJDIR)/main_simulator.o $(OUT)/libcore.a $(LIBDIR)/liblogger.so
	@mkdir -p $(BINDIR)
	$(CC) -o $@ $(OBJDIR)/main_simulator.o $(OUT)/libcore.a -L$(LIBDIR) -llogger \
	    -Wl,-rpath,'$$ORIGIN/../shared_libs' $(LDLIBS)

define PLUGIN_RULE
$(LIBDIR)/lib$(1).so: $(addprefix $(OBJDIR)/,$(PLUG_$(1):.c=.o)) $(OUT)/libcore.a $(LIBDIR)/liblogger.so
	@mkdir -p $(LIBDIR)
	$(CC) -shared -o $$@ $(addprefix $(OBJDIR)/,$(PLUG_$(1):.c=.o)) $(OUT)/libcore.a \
	    -L$(LIBDIR) -llogger -Wl,-rpath,'$$$$ORIGIN' $(LDLIBS)
endef
$(foreach p,$(PLUGINS),$(eval $(call PLUGIN_RULE,$(p))))

clean:
	rm -rf $(OUT)

.PHONY: all clean
