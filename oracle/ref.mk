# Builds the reference C11 library (Release flags: -O3 -DNDEBUG, as its CMake
# Release configuration) straight from its sources under
# /root/reference/native (read-only) into oracle/_ref/ for use as a checker.
# No reference file is copied; only the shared object lands in oracle/_ref/.
# Usage: make -f oracle/ref.mk [-j8]
REF  ?= /root/reference/native
OUT  ?= oracle/_ref
CC   ?= gcc
SRCS := version.c common/error.c common/alloc.c common/color_convert.c \
        common/planar_convert.c common/cpudetect.c common/crc32.c common/tpool.c \
        transform/transform.c transform/transform16.c transform/sparse_inverse.c \
        transform/quant.c transform/fastdiv.c transform/plane.c transform/base_scale.c \
        transform/pyramid_transform.c bitstream/bitio.c bitstream/frame_header.c \
        bitstream/slice_map.c bitstream/packet.c bitstream/band_tokens.c \
        bitstream/band_codec.c bitstream/layer_directory.c bitstream/v7_band_decode.c \
        bitstream/v7_frame_codec.c bitstream/slice_codec.c entropy/rice.c entropy/vlc.c \
        entropy/rans.c entropy/range.c entropy/block_coding.c codec/codec.c \
        codec/ctx_ceiling.c codec/base_encoder.c codec/base_residual.c \
        codec/base_decoder.c codec/base_policy.c codec/deblock.c codec/v7_scalable.c \
        codec/intra.c codec/decoder_ctx.c codec/gop_context.c mov/mov.c \
        image/image_container.c image/image_reader.c image/image_writer.c \
        image/half_map.c simd/dispatch.c simd/transform_avx2.c \
        simd/color_convert_avx2.c simd/planar_convert_avx2.c simd/transform_neon.c
OBJS := $(patsubst %.c,$(OUT)/obj/%.o,$(SRCS))
CFLAGS := -O3 -DNDEBUG -std=c11 -fPIC -I$(REF)/include -I$(REF)/src \
          -DTOPOS_GIT_COMMIT=\"reference\" -DTOPOS_BUILD_TARGET=\"x86_64-gcc\" -D_GNU_SOURCE

all: $(OUT)/libtopos_ref.so

$(OUT)/obj/common/color_convert.o: CFLAGS += -ffp-contract=off

$(OUT)/obj/%.o: $(REF)/src/%.c
	@mkdir -p $(dir $@)
	$(CC) $(CFLAGS) -c $< -o $@

$(OUT)/libtopos_ref.so: $(OBJS)
	$(CC) -shared -o $@ $^ -lpthread -lm

clean:
	rm -rf $(OUT)/obj $(OUT)/libtopos_ref.so
.PHONY: all clean
