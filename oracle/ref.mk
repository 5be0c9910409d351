# Builds the reference ompVC program from its own sources, where they lie under
# /root/reference/ompVC, into oracle/_ref/ (git-ignored).  Used only as a checker:
# golden vectors under tests/golden/ are written by running it in the authoring
# container.  Flags follow the reference makefile (ompVC/makefile:8-10).
#   make -f oracle/ref.mk            (from the repo root)
REFDIR ?= /root/reference/ompVC
OUT    := oracle/_ref
SRC    := exposuretime.c ftle.c globals.c integration.c io.c main.c memory.c mesh.c \
          mymath.c parameters.c residencetime.c strainrate.c tracers.c velocity.c velout.c
OBJ    := $(addprefix $(OUT)/obj/,$(SRC:.c=.o))

all: $(OUT)/ompflowVC

$(OUT)/obj/%.o: $(REFDIR)/%.c
	@mkdir -p $(OUT)/obj
	gcc -O3 -fopenmp -w -c $< -o $@

$(OUT)/ompflowVC: $(OBJ)
	gcc $(OBJ) -lm -fopenmp -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
