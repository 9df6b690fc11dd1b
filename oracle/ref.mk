# Builds the reference's CPU tools from their own sources, where they lie under
# /root/reference, into oracle/_ref/.  Only the tools whose sources compile on
# their own are built:
#   bppivot        baselines/MaximalBiPlex (apps/main.cpp + src/biplex/biplex.cpp)
#   graph_convert  preprocess/graph_conver.cpp
# iTraversal (baselines/SGMD22-MBPE) needs nlohmann/json, which is not installed,
# and gmkbpe itself needs CUDA/CUB plus headers missing from the tree; neither is
# built.  Usage:  make -f oracle/ref.mk  (from the repository root)
REF ?= /root/reference
OUT ?= oracle/_ref
CXX ?= g++

all: $(OUT)/bppivot $(OUT)/graph_convert

$(OUT):
	mkdir -p $(OUT)

$(OUT)/bppivot: $(REF)/baselines/MaximalBiPlex/apps/main.cpp $(REF)/baselines/MaximalBiPlex/src/biplex/biplex.cpp | $(OUT)
	$(CXX) -std=c++17 -O3 -w -I$(REF)/baselines/MaximalBiPlex/src/utils -I$(REF)/baselines/MaximalBiPlex/src/biplex \
	  -o $@ $(REF)/baselines/MaximalBiPlex/apps/main.cpp $(REF)/baselines/MaximalBiPlex/src/biplex/biplex.cpp

$(OUT)/graph_convert: $(REF)/preprocess/graph_conver.cpp | $(OUT)
	$(CXX) -std=c++17 -O3 -w -fopenmp -o $@ $(REF)/preprocess/graph_conver.cpp

clean:
	rm -rf $(OUT)

.PHONY: all clean
