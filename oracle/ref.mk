# Builds the reference primecount (library + CLI) from its sources where they
# lie under /root/reference, with g++ directly (no reference build system).
# Outputs go only into oracle/_ref/.  The build is test infrastructure: it is
# the checker for fixtures and the CPU baseline, never part of the product.
#
#   make -f oracle/ref.mk -j8
#
REF      ?= /root/reference
OUT      ?= oracle/_ref
CXX      ?= g++
CXXFLAGS ?= -O3 -std=gnu++17 -fopenmp -fPIC -DNDEBUG
DEFS_PC  := -DENABLE_LIBDIVIDE -DENABLE_MULTIARCH_x86_POPCNT -DENABLE_MULTIARCH_AVX512_VPOPCNT
DEFS_PS  := -DENABLE_MULTIARCH_x86_POPCNT -DENABLE_MULTIARCH_AVX512_BW -DENABLE_MULTIARCH_AVX512_VBMI2
INC_PC   := -I$(REF)/include -I$(REF)/src -I$(REF)/lib/primesieve/include
INC_PS   := -I$(REF)/lib/primesieve/include -I$(REF)/lib/primesieve/src

PC_SRC := api.cpp api_c.cpp BitSieve240.cpp FactorTable.cpp RiemannR.cpp P2.cpp P3.cpp \
  PhiTiny.cpp PiTable.cpp S1.cpp sieve/Sieve.cpp sieve/pre_sieve.cpp LoadBalancerP2.cpp \
  LoadBalancerS2.cpp LogarithmicIntegral.cpp StatusS2.cpp generate_primes.cpp nth_prime.cpp \
  nth_prime_sieve.cpp phi.cpp phi_vector.cpp pi_legendre.cpp pi_lehmer.cpp pi_meissel.cpp \
  pi_primesieve.cpp print.cpp util.cpp lmo/pi_lmo1.cpp lmo/pi_lmo2.cpp lmo/pi_lmo3.cpp \
  lmo/pi_lmo4.cpp lmo/pi_lmo5.cpp lmo/pi_lmo_parallel.cpp deleglise-rivat/S2_easy.cpp \
  deleglise-rivat/S2_hard.cpp deleglise-rivat/S2_trivial.cpp \
  deleglise-rivat/pi_deleglise_rivat.cpp gourdon/pi_gourdon.cpp gourdon/Phi0.cpp \
  gourdon/AC.cpp gourdon/B.cpp gourdon/D.cpp gourdon/LoadBalancerAC.cpp \
  gourdon/SegmentedPiTable.cpp gourdon/Sigma.cpp arch/x86/cpuid.cpp
APP_SRC := app/CmdOptions.cpp app/main.cpp app/help.cpp app/test.cpp
PS_SRC := api-c.cpp api.cpp CountPrintPrimes.cpp CpuInfo.cpp Erat.cpp EratSmall.cpp \
  EratMedium.cpp EratBig.cpp iterator-c.cpp iterator.cpp IteratorHelper.cpp LookupTables.cpp \
  MemoryPool.cpp PrimeGenerator.cpp nthPrime.cpp ParallelSieve.cpp popcount.cpp PreSieve.cpp \
  PrimeSieveClass.cpp RiemannR.cpp SievingPrimes.cpp arch/x86/cpuid.cpp

PC_OBJ  := $(addprefix $(OUT)/obj/pc/,$(PC_SRC:.cpp=.o))
APP_OBJ := $(addprefix $(OUT)/obj/app/,$(APP_SRC:.cpp=.o))
PS_OBJ  := $(addprefix $(OUT)/obj/ps/,$(PS_SRC:.cpp=.o))

all: $(OUT)/primecount $(OUT)/libprimecount_ref.so

$(OUT)/obj/pc/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(DEFS_PC) $(INC_PC) -c $< -o $@

$(OUT)/obj/app/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(DEFS_PC) $(INC_PC) -c $< -o $@

$(OUT)/obj/ps/%.o: $(REF)/lib/primesieve/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(DEFS_PS) $(INC_PS) -c $< -o $@

$(OUT)/libprimecount_ref.so: $(PC_OBJ) $(PS_OBJ)
	$(CXX) -shared -fopenmp -o $@ $^

$(OUT)/primecount: $(APP_OBJ) $(PC_OBJ) $(PS_OBJ)
	$(CXX) -fopenmp -o $@ $^

.PHONY: all
