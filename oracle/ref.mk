# Builds the reference ompVC program from its own sources, where they lie under
# /root/reference/ompVC, into oracle/_ref/ (git-ignored), and the reference
# mesh renumbering tool (cudaVC/ZOrderSort/main.cpp, one C++ file) beside it.
# Used only as checkers: golden vectors under tests/golden/ are written by
# running them in the authoring container.  Flags follow the reference makefile
# (ompVC/makefile:8-10).
#   make -f oracle/ref.mk            (from the repo root)
REFDIR ?= /root/reference/ompVC
ZODIR  ?= /root/reference/cudaVC/ZOrderSort
OUT    := oracle/_ref
SRC    := exposuretime.c ftle.c globals.c integration.c io.c main.c memory.c mesh.c \
          mymath.c parameters.c residencetime.c strainrate.c tracers.c velocity.c velout.c
OBJ    := $(addprefix $(OUT)/obj/,$(SRC:.c=.o))

all: $(OUT)/ompflowVC $(OUT)/zorder_sort

$(OUT)/obj/%.o: $(REFDIR)/%.c
	@mkdir -p $(OUT)/obj
	gcc -O3 -fopenmp -w -c $< -o $@

$(OUT)/ompflowVC: $(OBJ)
	gcc $(OBJ) -lm -fopenmp -o $@

$(OUT)/zorder_sort: $(ZODIR)/main.cpp
	@mkdir -p $(OUT)
	g++ -O2 -w $< -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
