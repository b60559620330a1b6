# oracle/_ref: the reference's own SPARSKIT and band-solver routines
# (ILUT, ILU(0), LUSOL, BiCGStab with its reverse-communication driver, AMUX,
# PENTA_FACT/PENTA_SOLVE/TRIA_MULTI), compiled from their sources where they
# lie in the reference tree with AMD flang, into one shared library used
# only by tests/test_oracle_ref.py to check the C restatement in this
# directory; and the reference's SHY file reader (io_shy.f90 with
# to_linear.f90) linked with shy_dump.f90 from this directory into
# _ref/shy_dump, used only by tests/test_shy_cpu.py to read the files the
# library writes; and the reference's debug-file checker (check_shympi_debug
# with clo.f90) into _ref/check_dbg, used only by the .dbg tests; and the
# reference's basin reader (basin.f90 with file.f90) linked with
# bas_dump.f90 from this directory into _ref/bas_dump, used only by
# tests/test_bas_cpu.py; and the reference's set_ev, element colouring,
# quad-tree element search and femintp_xi (evgeom.f90, subset.f90,
# quad_tree.f90 with what they use) linked with geom_ref.f90 from this
# directory into _ref/geom_ref, used only by tests/test_oracle_geom.py.
# Nothing is copied; the build writes only to oracle/_ref/.
#   make -f ref.mk            (needs the reference tree at REF)
REF   ?= /root/reference
FLANG ?= /opt/rocm/llvm/bin/flang
SPK   := $(REF)/src/external/sparskit
SRC   := $(SPK)/spk_ilut.f90 $(SPK)/spk_iters.f90 $(SPK)/spk_itaux.f90 $(SPK)/spk_blas.f90 \
         $(SPK)/spk_matvec.f90 $(SPK)/penta.f90
OUT   := _ref

SHYSRC := $(REF)/src/utils/generic/to_linear.f90 $(REF)/src/utils/shyutil/io_shy.f90

DBGSRC := $(REF)/src/utils/generic/clo.f90 $(REF)/src/utils/shympi/check_shympi_debug.f90

BASSRC := $(REF)/src/utils/generic/file.f90 $(REF)/src/utils/shyutil/basin.f90

# set_ev, the colouring, the quad tree and femintp_xi with what they use
GEOSRC := $(REF)/src/utils/generic/file.f90 $(REF)/src/utils/shyutil/basin.f90 \
          $(REF)/src/utils/shyutil/geom.f90 $(REF)/src/utils/shyutil/evgeom.f90 \
          $(REF)/src/utils/generic/stack.f90 $(REF)/src/utils/shyutil/quad_tree.f90 \
          $(REF)/src/utils/shyutil/find.f90 $(REF)/src/utils/shyutil/regular.f90 \
          $(REF)/src/utils/generic/round.f90 $(REF)/src/shyfem/subset.f90

all: $(OUT)/libspkref.so $(OUT)/shy_dump $(OUT)/check_dbg $(OUT)/bas_dump $(OUT)/geom_ref

# the reference's geometry, colouring and upwind-point search
$(OUT)/geom_ref: $(GEOSRC) geom_ref.f90
	@mkdir -p $(OUT)/geo
	cd $(OUT)/geo && for f in $(GEOSRC); do $(FLANG) -O1 -c $$f || exit 1; done
	$(FLANG) -O1 -module-dir $(OUT)/geo -o $@ geom_ref.f90 $(addprefix $(OUT)/geo/,$(notdir $(GEOSRC:.f90=.o)))

# the reference's basin (.bas) reader
$(OUT)/bas_dump: $(BASSRC) bas_dump.f90
	@mkdir -p $(OUT)/bas
	cd $(OUT)/bas && for f in $(BASSRC); do $(FLANG) -O1 -c $$f || exit 1; done
	$(FLANG) -O1 -module-dir $(OUT)/bas -o $@ bas_dump.f90 $(addprefix $(OUT)/bas/,$(notdir $(BASSRC:.f90=.o)))

# the reference's debug-file checker (lists one file, compares two)
$(OUT)/check_dbg: $(DBGSRC)
	@mkdir -p $(OUT)/dbg
	cd $(OUT)/dbg && for f in $(DBGSRC); do $(FLANG) -O1 -c $$f || exit 1; done
	$(FLANG) -O1 -o $@ $(addprefix $(OUT)/dbg/,$(notdir $(DBGSRC:.f90=.o)))

$(OUT)/shy_dump: $(SHYSRC) shy_dump.f90
	@mkdir -p $(OUT)/shy
	cd $(OUT)/shy && for f in $(SHYSRC); do $(FLANG) -O1 -c $$f || exit 1; done
	$(FLANG) -O1 -module-dir $(OUT)/shy -o $@ shy_dump.f90 $(addprefix $(OUT)/shy/,$(notdir $(SHYSRC:.f90=.o)))

$(OUT)/libspkref.so: $(SRC)
	@mkdir -p $(OUT)
	cd $(OUT) && for f in $(SRC); do $(FLANG) -O2 -fPIC -c $$f || exit 1; done
	$(FLANG) -shared -fPIC -o $@ $(addprefix $(OUT)/,$(notdir $(SRC:.f90=.o)))

clean:
	rm -rf $(OUT)

.PHONY: all clean
