# oracle/_ref: the reference's own SPARSKIT and band-solver routines
# (ILUT, ILU(0), LUSOL, BiCGStab with its reverse-communication driver, AMUX,
# PENTA_FACT/PENTA_SOLVE/TRIA_MULTI), compiled from their sources where they
# lie in the reference tree with AMD flang, into one shared library used
# only by tests/test_oracle_ref.py to check the C restatement in this
# directory.  Nothing is copied; the build writes only to oracle/_ref/.
#   make -f ref.mk            (needs the reference tree at REF)
REF   ?= /root/reference
FLANG ?= /opt/rocm/llvm/bin/flang
SPK   := $(REF)/src/external/sparskit
SRC   := $(SPK)/spk_ilut.f90 $(SPK)/spk_iters.f90 $(SPK)/spk_itaux.f90 $(SPK)/spk_blas.f90 \
         $(SPK)/spk_matvec.f90 $(SPK)/penta.f90
OUT   := _ref

all: $(OUT)/libspkref.so

$(OUT)/libspkref.so: $(SRC)
	@mkdir -p $(OUT)
	cd $(OUT) && for f in $(SRC); do $(FLANG) -O2 -fPIC -c $$f || exit 1; done
	$(FLANG) -shared -fPIC -o $@ $(addprefix $(OUT)/,$(notdir $(SRC:.f90=.o)))

clean:
	rm -rf $(OUT)

.PHONY: all clean
